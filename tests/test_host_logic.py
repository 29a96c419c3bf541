"""Host-side logic of the Python layer (CPU only)."""

import numpy as np
import pytest

from finitevolumegpu_amd.Simulator import BaseSimulator, BoundaryCondition


def test_boundary_condition_packing():
    T = BoundaryCondition.Type
    bc = BoundaryCondition({"north": T.Reflective, "south": T.Periodic, "east": T.Dirichlet, "west": T.Outflow})
    code = int(bc.asCodedInt())
    assert code == (3 << 24) | (2 << 16) | (0 << 8) | 4
    back = BoundaryCondition(BoundaryCondition.getTypes(code))
    assert (back.north, back.south, back.east, back.west) == (T.Reflective, T.Periodic, T.Dirichlet, T.Outflow)
    assert int(BoundaryCondition().asCodedInt()) == 0x03030303


def test_neumann_rejected():
    T = BoundaryCondition.Type
    with pytest.raises(NotImplementedError):
        BoundaryCondition({"north": T.Neumann, "south": T.Reflective, "east": T.Reflective, "west": T.Reflective})


class _FakeSim(BaseSimulator):
    """BaseSimulator without a device: substeps only record their dt."""

    def __init__(self, cfl):
        self.t = 0.0
        self.nt = 0
        self.dt = None
        self.cfl_scale = 0.9
        self.num_substeps = 2
        self.cfl = cfl
        self.substeps = []
        self.dt_calls = []
        import logging
        self.logger = logging.getLogger("fake")

    def substep(self, dt, step_number):
        self.substeps.append((float(dt), step_number))

    def computeDt(self):
        self.dt_calls.append(self.nt)
        return float(np.float32(self.cfl)) * 0.5

    def check(self):
        pass


def test_simulate_dt_policy():
    sim = _FakeSim(0.1)
    t_end = 0.045 * 150.25
    sim.simulate(t_end)
    assert sim.dt_calls[:2] == [0, 100]
    assert sim.nt in (151, 152)
    assert abs(sim.t - t_end) < 1e-6
    # every step passes a float32 dt to both substeps, numbered 0 and 1
    assert [s for _, s in sim.substeps[:4]] == [0, 1, 0, 1]
    assert all(np.float32(d) == d for d, _ in sim.substeps)
    assert isinstance(sim.t, float)


def test_simulate_fixed_dt():
    sim = _FakeSim(0.1)
    sim.simulate(1.0, dt=0.3)
    assert sim.dt_calls == []
    assert sim.nt == 4
    assert abs(sim.substeps[-1][0] - np.float32(0.1)) < 1e-7
