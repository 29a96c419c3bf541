"""Subdomain group with on-device ghost-cell exchange (GPUSimulators/SHMEMSimulator.py,
SHMEMSimulatorGroup.py).

The reference runs one simulator per subdomain, each in its own context (all
on device 0 in its single-GPU configuration, SHMEMSimulatorGroup.py:56-67), and
before every substep copies the ghost cells between neighbours through host
memory: north/south interior rows first, then east/west columns including the
ghost rows just filled (SHMEMSimulator.py:200-263).  Here the subdomains live
on one GPU and the exchange stays in HBM: the regions of each leg are copied
device to device by one batched launch (``fvmi_copy_regions_f32``) on an
exchange stream that waits for every subdomain's last substep, and every
subdomain's stream waits for the exchange before its next substep.

``SHMEMGrid`` keeps the reference's grid shape rule (SHMEMSimulatorGroup.py:97-156):
the subdomain count is split into factors of least sum, searched over
divisors d with d*d < n, a lone 4 is split into 2 x 2, the list is padded with
ones and sorted in descending order (so 9 subdomains give a 9 x 1 grid, as in
the reference).
"""

import ctypes
import functools
import logging

import numpy as np

from . import _lib
from .HipContext import Event, HipContext, Stream, device_count
from .Simulator import BaseSimulator, BoundaryCondition


def _least_sum_factors(n, parts):
    """(sum, factors) of the split of ``n`` into at most ``parts`` factors with the
    least sum, trying each divisor d (2 <= d, d*d < n) in increasing order and
    keeping the first split of a strictly smaller sum."""

    @functools.lru_cache(maxsize=None)
    def best(m, k):
        if k == 1:
            return m, (m,)
        cost, split = m, (m,)
        d = 2
        while d * d < m:
            if m % d == 0:
                sub_cost, sub = best(m // d, k - 1)
                if sub_cost + d < cost:
                    cost, split = sub_cost + d, (d,) + sub
            d += 1
        return cost, split

    return best(int(n), int(parts))


class SHMEMGrid:
    """Grid of subdomains, one context each (SHMEMSimulatorGroup.py:31-96)."""

    def __init__(self, ngpus=None, ndims=2):
        self.logger = logging.getLogger(__name__)
        if ndims != 2:
            raise ValueError("Unsupported number of dimensions. Must be two at the moment")
        if ngpus is None:
            ngpus = device_count()
        self.ngpus = int(ngpus)
        self.ndims = ndims
        self.grid = SHMEMGrid.getGrid(self.ngpus, self.ndims)
        # one context per subdomain, all on device 0 (the reference's one-GPU set-up)
        self.cuda_contexts = [HipContext(device=0, autotuning=False) for _ in range(self.ngpus)]

    def getCoordinate(self, index):
        return index % self.grid[0], index // self.grid[0]

    def getIndex(self, i, j):
        return int(j * self.grid[0] + i)

    def getEast(self, index):
        i, j = self.getCoordinate(index)
        return self.getIndex((i + 1) % self.grid[0], j)

    def getWest(self, index):
        i, j = self.getCoordinate(index)
        return self.getIndex((i + self.grid[0] - 1) % self.grid[0], j)

    def getNorth(self, index):
        i, j = self.getCoordinate(index)
        return self.getIndex(i, (j + 1) % self.grid[1])

    def getSouth(self, index):
        i, j = self.getCoordinate(index)
        return self.getIndex(i, (j + self.grid[1] - 1) % self.grid[1])

    @staticmethod
    def getGrid(num_gpus, num_dims):
        if not isinstance(num_gpus, (int, np.integer)) or not isinstance(num_dims, (int, np.integer)):
            raise TypeError("num_gpus and num_dims must be integers")
        factors = list(_least_sum_factors(num_gpus, num_dims)[1])
        if len(factors) < num_dims:
            if 4 in factors:
                factors.remove(4)
                factors += [2, 2]
            factors += [1] * (num_dims - len(factors))
        return np.sort(np.array(factors))[::-1]


def exchange_extents(nx, ny, gx, gy):
    """[x0, y0, width, height] of the regions of a padded (ny + 2 gy) x (nx + 2 gx)
    subdomain that the exchange reads (interior cells next to each edge) and
    writes (the ghost cells beyond it), in the order east, west, north, south
    (SHMEMSimulator.py:124-136).  East/west regions span the full padded height,
    so they carry the corners filled by the north/south leg."""
    read_e = np.array([nx, 0, gx, ny + 2 * gy])
    read_w = np.array([gx, 0, gx, ny + 2 * gy])
    read_n = np.array([gx, ny, nx, gy])
    read_s = np.array([gx, gy, nx, gy])
    write_e = read_e + np.array([gx, 0, 0, 0])
    write_w = read_w - np.array([gx, 0, 0, 0])
    write_n = read_n + np.array([0, gy, 0, 0])
    write_s = read_s - np.array([0, gy, 0, 0])
    return {"read": (read_e, read_w, read_n, read_s), "write": (write_e, write_w, write_n, write_s)}


class SHMEMSimulator(BaseSimulator):
    """Steps a group of subdomain simulators as one (SHMEMSimulator.py:31-263).

    ``sims[i]`` is the subdomain at ``grid.getCoordinate(i)``; all must use the
    same scheme and grid size.  The global boundary conditions are those of
    ``sims[0]``: edges between subdomains (and periodic outer edges, which wrap
    to the opposite subdomain) become Dirichlet and are filled by the exchange.
    """

    def __init__(self, sims, grid):
        self.logger = logging.getLogger(__name__)
        if len(sims) < 2:
            raise ValueError("SHMEMSimulator needs at least two subdomains")
        if len(sims) != int(np.prod(grid.grid)):
            raise ValueError("%d simulators for a %s grid" % (len(sims), list(grid.grid)))
        self.sims = sims
        self.grid = grid
        s0 = sims[0]
        for s in sims:
            if (s.nx, s.ny, len(s.getOutput().gpu_variables)) != (s0.nx, s0.ny, len(s0.getOutput().gpu_variables)):
                raise ValueError("all subdomains must have the same size and variables")
        # BaseSimulator state mirrors the first subdomain (SHMEMSimulator.py:50-60)
        self.context = s0.context
        self.nx, self.ny, self.dx, self.dy = s0.nx, s0.ny, s0.dx, s0.dy
        self.cfl_scale = s0.cfl_scale
        self.num_substeps = s0.num_substeps
        self.block_size = s0.block_size
        self.stream = s0.stream
        self.internal_stream = s0.internal_stream
        self.t = 0.0
        self.nt = 0
        self.dt = None
        bc = s0.getBoundaryConditions()
        self.boundary_conditions = s0.boundary_conditions
        T = BoundaryCondition.Type
        n = len(sims)
        self.east, self.west, self.north, self.south = [None] * n, [None] * n, [None] * n, [None] * n
        for i, sim in enumerate(sims):
            self.east[i], self.west[i] = grid.getEast(i), grid.getWest(i)
            self.north[i], self.south[i] = grid.getNorth(i), grid.getSouth(i)
            new_bc = {"north": T.Dirichlet, "south": T.Dirichlet, "east": T.Dirichlet, "west": T.Dirichlet}
            gi, gj = grid.getCoordinate(i)
            # outer edges keep the global condition unless it is periodic
            # (SHMEMSimulator.py:101-117, per subdomain)
            if gi == 0 and bc.west != T.Periodic:
                self.west[i] = None
                new_bc["west"] = bc.west
            if gj == 0 and bc.south != T.Periodic:
                self.south[i] = None
                new_bc["south"] = bc.south
            if gi == grid.grid[0] - 1 and bc.east != T.Periodic:
                self.east[i] = None
                new_bc["east"] = bc.east
            if gj == grid.grid[1] - 1 and bc.north != T.Periodic:
                self.north[i] = None
                new_bc["north"] = bc.north
            sim.setBoundaryConditions(BoundaryCondition(new_bc))
        self.nvars = len(s0.getOutput().gpu_variables)
        a0 = s0.getOutput()[0]
        self.gc_x, self.gc_y = int(a0.x_halo), int(a0.y_halo)
        ext = exchange_extents(int(self.nx), int(self.ny), self.gc_x, self.gc_y)
        self.read_e, self.read_w, self.read_n, self.read_s = ext["read"]
        self.write_e, self.write_w, self.write_n, self.write_s = ext["write"]
        self.exchange_stream = Stream()
        self._done = [Event() for _ in sims]
        self._exchanged = Event()
        self.logger.debug("Initialized %d subdomains", n)

    # -- exchange -----------------------------------------------------------

    @staticmethod
    def _region(src_arr, src_ext, dst_arr, dst_ext):
        x0, y0, w, h = (int(v) for v in src_ext)
        u0, v0 = int(dst_ext[0]), int(dst_ext[1])
        for arr, x, y in ((src_arr, x0, y0), (dst_arr, u0, v0)):
            if x < 0 or y < 0 or x + w > arr.shape[1] or y + h > arr.shape[0]:
                raise ValueError("exchange region outside the padded buffer")
        r = _lib.RegionCopy()
        r.src = src_arr.row_ptr(y0, x0)
        r.dst = dst_arr.row_ptr(v0, u0)
        r.src_pitch, r.dst_pitch, r.width, r.height = src_arr.pitch, dst_arr.pitch, w, h
        return r

    def _leg(self, pairs):
        """One batched copy of (neighbour, read extent, subdomain, write extent) regions."""
        regions = []
        for src, rd, dst, wr in pairs:
            for k in range(self.nvars):
                regions.append(self._region(self.sims[src].u0[k], rd, self.sims[dst].u0[k], wr))
        if regions:
            table = (_lib.RegionCopy * len(regions))(*regions)
            _lib.call("fvmi_copy_regions_f32", table, len(regions), self.exchange_stream.handle)

    def exchange(self):
        """Ghost cells of every subdomain's current state from its neighbours:
        north/south interior rows, then east/west columns over the full height
        (SHMEMSimulator.py:200-217)."""
        for i, sim in enumerate(self.sims):
            self._done[i].record(sim.stream)
            self.exchange_stream.wait_for_event(self._done[i])
        ns, ew = [], []
        for i in range(len(self.sims)):
            if self.north[i] is not None:
                ns.append((self.north[i], self.read_s, i, self.write_n))
            if self.south[i] is not None:
                ns.append((self.south[i], self.read_n, i, self.write_s))
        for i in range(len(self.sims)):
            if self.east[i] is not None:
                ew.append((self.east[i], self.read_w, i, self.write_e))
            if self.west[i] is not None:
                ew.append((self.west[i], self.read_e, i, self.write_w))
        self._leg(ns)
        self._leg(ew)
        self._exchanged.record(self.exchange_stream)
        for sim in self.sims:
            sim.stream.wait_for_event(self._exchanged)

    # -- BaseSimulator interface --------------------------------------------

    def substep(self, dt, step_number):
        self.exchange()
        for sim in self.sims:
            sim.substep(dt, step_number)

    def getOutput(self):
        # the reference returns the first subdomain's state here (SHMEMSimulator.py:157-160),
        # and BaseSimulator.download therefore the first subdomain's fields
        return self.sims[0].getOutput()

    def download_all(self, variables=None):
        """The whole domain: the subdomains' interiors assembled by grid position
        (no counterpart in the reference, whose download returns subdomain 0)."""
        parts = [sim.download(variables) for sim in self.sims]
        nx, ny = int(self.nx), int(self.ny)
        gx, gy = int(self.grid.grid[0]), int(self.grid.grid[1])
        out = []
        for k in range(len(parts[0])):
            full = np.empty((gy * ny, gx * nx), dtype=parts[0][k].dtype)
            for i, p in enumerate(parts):
                ci, cj = self.grid.getCoordinate(i)
                full[cj * ny:(cj + 1) * ny, ci * nx:(ci + 1) * nx] = p[k]
            out.append(full)
        return out

    def synchronize(self):
        for sim in self.sims:
            sim.synchronize()
        self.exchange_stream.synchronize()

    def check(self):
        for sim in self.sims:
            sim.check()

    def computeDt(self):
        """Minimum over the subdomains (SHMEMSimulator.py:171-184)."""
        self.synchronize()
        return min(sim.computeDt() for sim in self.sims)

    def getExtent(self, index=0):
        width = self.sims[index].nx * self.sims[index].dx
        height = self.sims[index].ny * self.sims[index].dy
        i, j = self.grid.getCoordinate(index)
        x0, y0 = i * width, j * height
        return [x0, x0 + width, y0, y0 + height]
