"""runSimulation on the GPU with the reference's own drivers' set-ups
(singleGPUTesting.py: EE2D_KP07_dimsplit on the Kelvin-Helmholtz case;
mpiTesting.py: the same inside an MPISimulator with a fixed dt): every saved
frame equals the least_significant_digit=3 rounding of the state that a second
simulator reaches through the same simulate() calls."""

import numpy as np
import pytest
from scipy.io import netcdf_file

from finitevolumegpu_amd import Common

pytestmark = pytest.mark.gpu


def _kh_args(ctx, nx, ny):
    from finitevolumegpu_amd.helpers import InitialConditions as IC
    np.random.seed(42)
    args = IC.genKelvinHelmholtz(nx, ny, 1.4)
    args["context"] = ctx
    args["theta"] = 1.2
    return args


def _replay(gen, args, times, dt, names):
    sim = gen(**args)
    steps = np.concatenate([[times[0]], np.diff(times)])
    frames = []
    for t_step in steps:
        if t_step > 0.0:
            sim.simulate(t_step, dt)
        frames.append([Common.quantize(a, 3) for a in sim.download(list(range(len(names))))])
    return frames


@pytest.mark.parametrize("wrap", [False, True])
def test_run_simulation_kelvin_helmholtz(ctx, tmp_path, wrap):
    from finitevolumegpu_amd.EE2D_KP07_dimsplit import EE2D_KP07_dimsplit
    from finitevolumegpu_amd.MPISimulator import MPISimulator
    print("case runSimulation KH, MPISimulator" if wrap else "case runSimulation KH", flush=True)
    nx, ny = 96, 48

    def gen(**kw):
        sim = EE2D_KP07_dimsplit(**kw)
        return MPISimulator(sim) if wrap else sim

    args = _kh_args(ctx, nx, ny)
    names = ["rho", "rho_u", "rho_v", "E"]
    times = np.linspace(0, 0.05, 4)
    dt = 0.0005 if wrap else None
    out, prof, mpi = Common.runSimulation(gen, args, str(tmp_path / "kh.nc"), times, names, dt)
    assert (mpi is not None) == wrap
    want = _replay(gen, args, times, dt, names)
    f = netcdf_file(out, "r", mmap=False)
    assert f.dimensions == {"time": 4, "x": nx, "y": ny}
    for k in range(len(times)):
        for i, name in enumerate(names):
            assert np.array_equal(f.variables[name][k], want[k][i]), (name, k)
    assert not np.array_equal(f.variables["rho"][0], f.variables["rho"][-1])  # the state evolved
    f.close()


def test_run_simulation_shallow_water(ctx, tmp_path):
    from finitevolumegpu_amd.helpers import InitialConditions as IC
    from finitevolumegpu_amd.KP07_dimsplit import KP07_dimsplit
    print("case runSimulation KP07_dimsplit", flush=True)
    h, hu, hv, dx, dy = IC.bump(64, 64, 100.0, 100.0, bump_size=20.0)
    args = {"context": ctx, "h0": h.astype(np.float32), "hu0": hu.astype(np.float32), "hv0": hv.astype(np.float32),
            "nx": 64, "ny": 64, "dx": dx, "dy": dy, "g": 9.81}
    times = np.linspace(0, 2.0, 3)
    names = ["h", None, "hv"]  # the None entry is skipped, hv keeps index 2
    out, _, _ = Common.runSimulation(KP07_dimsplit, args, str(tmp_path / "swe.nc"), times, names)
    sim = KP07_dimsplit(**args)
    f = netcdf_file(out, "r", mmap=False)
    assert set(f.variables) == {"time", "x", "y", "h", "hv"}
    for k, t_step in enumerate(np.concatenate([[times[0]], np.diff(times)])):
        if t_step > 0:
            sim.simulate(t_step)
        h_k, hv_k = sim.download([0, 2])
        assert np.array_equal(f.variables["h"][k], Common.quantize(h_k, 3))
        assert np.array_equal(f.variables["hv"][k], Common.quantize(hv_k, 3))
    f.close()
