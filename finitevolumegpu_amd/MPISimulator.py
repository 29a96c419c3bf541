"""Domain-decomposition wrapper reduced to one GPU (GPUSimulators/MPISimulator.py).

The reference splits the domain into row strips, one MPI rank per GPU, and
after every substep exchanges two ghost rows/columns with its neighbours
through host memory (MPISimulator.py:311-480), taking the global dt as an
Allreduce MIN.  On one GPU the rank is its own neighbour across every periodic
edge: ``MPISimulator`` keeps that structure -- the wrapped simulator's
periodic edges become Dirichlet ("use what is in memory", :247-267) and an
on-device halo exchange fills those ghost cells after every substep and once
at construction (:308) -- with no MPI, host staging or inter-GPU traffic.

``MPIGrid`` keeps the neighbour arithmetic of the reference (:34-148) for a
communicator of any size, so a multi-rank version only has to replace the
exchange; here ``comm`` defaults to a single rank.
"""

import ctypes
import logging
import time

import numpy as np

from . import _lib
from .Simulator import BaseSimulator, BoundaryCondition


class SingleRankComm:
    """The part of an MPI communicator that MPIGrid uses, for one rank."""
    rank = 0
    size = 1

    def allreduce_min(self, value):
        return value


class MPIGrid:
    """Row-strip process grid (MPISimulator.py:34-148; grid forced to [1, N])."""

    def __init__(self, comm=None, ndims=2):
        if ndims != 2:
            raise ValueError("Unsupported number of dimensions. Must be two at the moment")
        self.comm = comm if comm is not None else SingleRankComm()
        if self.comm.size < 1:
            raise ValueError("Must have at least one node")
        self.grid = np.array([1, self.comm.size])

    def getCoordinate(self, rank=None):
        if rank is None:
            rank = self.comm.rank
        return rank % self.grid[0], rank // self.grid[0]

    def getRank(self, i, j):
        return int(j * self.grid[0] + i)

    def getEast(self):
        i, j = self.getCoordinate()
        return self.getRank((i + 1) % self.grid[0], j)

    def getWest(self):
        i, j = self.getCoordinate()
        return self.getRank((i + self.grid[0] - 1) % self.grid[0], j)

    def getNorth(self):
        i, j = self.getCoordinate()
        return self.getRank(i, (j + 1) % self.grid[1])

    def getSouth(self):
        i, j = self.getCoordinate()
        return self.getRank(i, (j + self.grid[1] - 1) % self.grid[1])


class MPISimulator(BaseSimulator):
    """Wraps a dimensionally split simulator; see the module docstring."""

    def __init__(self, sim, grid=None):
        self.logger = logging.getLogger(__name__)
        self.profiling_data_mpi = {"start": {}, "end": {}, "n_time_steps": 0}
        for key in ("t_mpi_halo_exchange", "t_mpi_step"):
            self.profiling_data_mpi["start"][key] = 0.0
            self.profiling_data_mpi["end"][key] = 0.0
        self.sim = sim
        self.grid = grid if grid is not None else MPIGrid()
        if self.grid.comm.size != 1:
            raise NotImplementedError("this build runs one rank per process on one GPU")
        # mirror BaseSimulator state (the wrapped simulator holds the device state)
        self.context = sim.context
        self.nx, self.ny, self.dx, self.dy = sim.nx, sim.ny, sim.dx, sim.dy
        self.cfl_scale = sim.cfl_scale
        self.num_substeps = sim.num_substeps
        self.block_size = sim.block_size
        self.stream = sim.stream
        self.internal_stream = sim.internal_stream
        self.t = 0.0
        self.nt = 0
        self.dt = None

        bc = sim.getBoundaryConditions()
        T = BoundaryCondition.Type
        # neighbours across periodic edges are this rank itself (MPISimulator.py:247-267)
        self.east = self.grid.getEast() if bc.east == T.Periodic else None
        self.west = self.grid.getWest() if bc.west == T.Periodic else None
        self.north = self.grid.getNorth() if bc.north == T.Periodic else None
        self.south = self.grid.getSouth() if bc.south == T.Periodic else None
        if (self.east is None) != (self.west is None) or (self.north is None) != (self.south is None):
            raise ValueError("a periodic edge needs a periodic opposite edge")
        new_bc = BoundaryCondition({
            "north": T.Dirichlet if self.north is not None else bc.north,
            "south": T.Dirichlet if self.south is not None else bc.south,
            "east": T.Dirichlet if self.east is not None else bc.east,
            "west": T.Dirichlet if self.west is not None else bc.west,
        })
        sim.setBoundaryConditions(new_bc)
        self.boundary_conditions = sim.boundary_conditions
        self.nvars = len(sim.getOutput().gpu_variables)
        self.full_exchange()
        sim.synchronize()

    def full_exchange(self):
        """On-device exchange of the current state's ghost cells (MPISimulator.py:371-480)."""
        t0 = time.time()
        state = self.sim.getOutput()
        ptrs = (ctypes.c_void_p * self.nvars)(*[state[k].data_ptr for k in range(self.nvars)])
        pitches = (ctypes.c_int * self.nvars)(*[state[k].pitch for k in range(self.nvars)])
        _lib.call("fvmi_halo_exchange_periodic", int(self.nx), int(self.ny), self.nvars, ptrs, pitches,
                  int(self.east is not None), int(self.north is not None), self.sim.stream.handle)
        self.profiling_data_mpi["start"]["t_mpi_halo_exchange"] += t0
        self.profiling_data_mpi["end"]["t_mpi_halo_exchange"] += time.time()

    def substep(self, dt, step_number):
        t0 = time.time()
        self.sim.substep(dt, step_number)  # full-domain substep (buffers swapped inside)
        self.profiling_data_mpi["start"]["t_mpi_step"] += t0
        self.profiling_data_mpi["end"]["t_mpi_step"] += time.time()
        self.full_exchange()
        self.profiling_data_mpi["n_time_steps"] += 1

    def getOutput(self):
        return self.sim.getOutput()

    def download(self, variables=None):
        return self.sim.download(variables)

    def synchronize(self):
        self.sim.synchronize()

    def check(self):
        return self.sim.check()

    def computeDt(self):
        """Allreduce MIN over ranks of the local dt (MPISimulator.py:349-354)."""
        return self.grid.comm.allreduce_min(self.sim.computeDt())

    def getExtent(self):
        width = self.sim.nx * self.sim.dx
        height = self.sim.ny * self.sim.dy
        i, j = self.grid.getCoordinate()
        x0 = i * width
        y0 = j * height
        return [x0, x0 + width, y0, y0 + height]
