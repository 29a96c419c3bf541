"""SHMEMSimulator on one GPU: a domain split into subdomains with the on-device
ghost exchange gives the single-domain simulator's fields bitwise, for the
dimensionally split and the unsplit schemes, reflective, periodic and mixed
global boundaries, fixed dt and the simulate() dt policy."""

import numpy as np
import pytest

from oracle import fv_oracle as O
from tests._cases import make_gpu_sim, state

pytestmark = pytest.mark.gpu

R, P = O.REFLECTIVE, O.PERIODIC


def _unsplit(ctx, scheme, fields, nx, ny, bc, math):
    from finitevolumegpu_amd.HLL import HLL
    from finitevolumegpu_amd.KP07 import KP07
    from finitevolumegpu_amd.Simulator import BoundaryCondition
    b = BoundaryCondition(BoundaryCondition.getTypes(bc))
    if scheme == "kp07":
        return KP07(ctx, *fields, nx, ny, 1.0, 1.0, 9.81, boundary_conditions=b, math=math)
    return HLL(ctx, *fields, nx, ny, 1.0, 1.0, 9.81, boundary_conditions=b, math=math)


def _make(ctx, scheme, fields, nx, ny, bc, math):
    if scheme in ("kp07", "hll"):
        return _unsplit(ctx, scheme, fields, nx, ny, bc, math)
    return make_gpu_sim(ctx, scheme, fields, nx, ny, 1.0, 1.0, bc, math=math)


def _pair(ctx, scheme, n, nx, ny, bc, math, seed):
    from finitevolumegpu_amd.SHMEMSimulator import SHMEMGrid, SHMEMSimulator
    grid = SHMEMGrid(ngpus=n)
    gx, gy = int(grid.grid[0]), int(grid.grid[1])
    NX, NY = gx * nx, gy * ny
    fields = state("ee2d_kp07_dimsplit" if scheme == "ee2d_kp07_dimsplit" else "kp07_dimsplit", NX, NY, seed)
    single = _make(ctx, scheme, fields, NX, NY, bc, math)
    subs = []
    for i in range(n):
        ci, cj = grid.getCoordinate(i)
        part = [np.ascontiguousarray(f[cj * ny:(cj + 1) * ny, ci * nx:(ci + 1) * nx]) for f in fields]
        subs.append(_make(grid.cuda_contexts[i], scheme, part, nx, ny, bc, math))
    return single, SHMEMSimulator(subs, grid), grid


CASES = [
    ("kp07_dimsplit", 4, 40, 24, O.pack_bc(R, R, R, R), "exact"),
    ("kp07_dimsplit", 4, 40, 24, O.pack_bc(P, P, P, P), "fast"),
    ("kp07_dimsplit", 6, 21, 17, O.pack_bc(P, R, R, P), "exact"),
    ("hll2", 2, 33, 30, O.pack_bc(R, R, P, P), "fast"),
    ("waf", 3, 20, 26, O.pack_bc(P, P, R, R), "exact"),
    ("ee2d_kp07_dimsplit", 2, 32, 48, O.pack_bc(R, R, P, P), "exact"),
    ("hll", 4, 30, 20, O.pack_bc(R, P, P, R), "exact"),
    ("kp07", 4, 30, 20, O.pack_bc(P, P, P, P), "exact"),
]


@pytest.mark.parametrize("scheme,n,nx,ny,bc,math", CASES)
def test_subdomains_equal_single_domain(ctx, scheme, n, nx, ny, bc, math):
    print("case shmem", scheme, n, nx, ny, bc, math, flush=True)
    single, shm, grid = _pair(ctx, scheme, n, nx, ny, bc, math, seed=n)
    dt = np.float32(0.9 * single.computeDt()) if scheme != "ee2d_kp07_dimsplit" else np.float32(0.05)
    for _ in range(12):
        single.step(dt)
        shm.step(dt)
    want = single.download()
    got = shm.download_all()
    for k in range(len(want)):
        assert np.array_equal(got[k], want[k]), (k, float(np.abs(got[k] - want[k]).max()))
    # the reference's download is the first subdomain's
    first = shm.download()
    assert np.array_equal(first[0], got[0][:ny, :nx])


def test_simulate_takes_the_global_dt(ctx):
    print("case shmem simulate", flush=True)
    bc = O.pack_bc(R, R, R, R)
    single, shm, grid = _pair(ctx, "kp07_dimsplit", 4, 32, 32, bc, "fast", seed=7)
    single.simulate(6.0)
    shm.simulate(6.0)
    assert shm.simSteps() == single.simSteps() and shm.simTime() == single.simTime()
    assert shm.computeDt() == single.computeDt()
    for a, b in zip(shm.download_all(), single.download()):
        assert np.array_equal(a, b)


def test_subdomain_boundaries(ctx):
    from finitevolumegpu_amd.Simulator import BoundaryCondition
    T = BoundaryCondition.Type
    _, shm, grid = _pair(ctx, "kp07_dimsplit", 4, 16, 16, O.pack_bc(P, P, R, R), "exact", seed=1)
    # grid 2 x 2: north/south periodic (wrap to the other row), east/west reflective outside
    assert grid.grid.tolist() == [2, 2]
    for i, sim in enumerate(shm.sims):
        gi, gj = grid.getCoordinate(i)
        b = sim.getBoundaryConditions()
        assert b.north == T.Dirichlet and b.south == T.Dirichlet
        assert b.west == (T.Reflective if gi == 0 else T.Dirichlet)
        assert b.east == (T.Reflective if gi == 1 else T.Dirichlet)
        assert (shm.west[i] is None) == (gi == 0) and (shm.east[i] is None) == (gi == 1)
        assert shm.north[i] == grid.getNorth(i) and shm.south[i] == grid.getSouth(i)
