"""Simulation output (finitevolumegpu_amd/Common.py; the reference's
runSimulation / DataDumper, GPUSimulators/Common.py:92-208, 361-431) on the
CPU: the least_significant_digit rounding against its closed form, unique
output names, the netCDF-3 dataset read back with scipy's own reader, and
runSimulation driven by a host-only simulator."""

import json

import numpy as np
import pytest
from scipy.io import netcdf_file

from finitevolumegpu_amd import Common


@pytest.mark.parametrize("lsd,step", [(3, 1.0 / 1024), (1, 1.0 / 16), (0, 1.0), (2, 1.0 / 128), (-1, 8.0)])
def test_quantize_keeps_binary_digits(lsd, step):
    """bits = ceil(log2(10**lsd)) binary digits after the point: the values are
    the nearest multiples of 2**-bits (ties to even), the first power of two
    not coarser than 10**-lsd."""
    assert step <= 10.0 ** -lsd < 2 * step
    rng = np.random.default_rng(1)
    x = rng.uniform(-50, 50, 1000).astype(np.float32)
    q = Common.quantize(x, lsd)
    assert q.dtype == np.float32
    k = q.astype(np.float64) / step
    assert np.array_equal(k, np.round(k))  # multiples of the step
    assert np.all(np.abs(q.astype(np.float64) - x) <= step / 2)
    # nearest multiple, ties to even (numpy.around)
    assert np.array_equal(q, (np.around(x.astype(np.float64) / step) * step).astype(np.float32))


def test_quantize_examples():
    q = Common.quantize(np.float32([0.25431, -1.0006, 3.14159, 1.0 / 2048, 3.0 / 2048]), 3)
    # 260.41/1024, -1024.61/1024, 3216.99/1024, and the ties 0.5 -> 0, 1.5 -> 2
    assert q.tolist() == [260 / 1024, -1025 / 1024, 3217 / 1024, 0.0, 2 / 1024]


def test_datadumper_never_overwrites(tmp_path):
    path = tmp_path / "sub" / "out.nc"
    with Common.DataDumper(str(path), mode="w") as d:
        d.ncfile.createDimension("x", 2)
        assert d.filename == str(path)
    with Common.DataDumper(str(path), mode="w") as d:
        assert d.filename == str(tmp_path / "sub" / "out_0000.nc")
        d.ncfile.createDimension("x", 2)
    d = Common.DataDumper(str(path), mode="w")
    assert d.filename == str(tmp_path / "sub" / "out_0001.nc")
    assert Common.DataDumper(str(path)).filename == str(path)  # reading keeps the name


def test_netcdf3_dataset_round_trip(tmp_path):
    path = str(tmp_path / "a.nc")
    ds = Common.Netcdf3Dataset(path, "w")
    ds.title = "test"
    ds.createDimension("time", 2)
    ds.createDimension("x", 3)
    t = ds.createVariable("time", "f", "time")
    v = ds.createVariable("v", np.dtype("float32").char, ("time", "x"), zlib=True, least_significant_digit=3)
    t[:] = [0.0, 0.5]
    v[0, :] = [0.1, 0.2, 0.3]
    v[1, :] = [1.00049, -2.0, 7.77777]
    ds.close()
    with pytest.raises(FileExistsError):
        Common.Netcdf3Dataset(path, "w", clobber=False)
    f = netcdf_file(path, "r", mmap=False)
    assert f.title == b"test"
    assert f.dimensions == {"time": 2, "x": 3}
    got = f.variables["v"][:]
    want = Common.quantize(np.float32([[0.1, 0.2, 0.3], [1.00049, -2.0, 7.77777]]), 3)
    assert np.array_equal(got, want)
    assert int(f.variables["v"].least_significant_digit) == 3
    assert f.variables["time"][:].tolist() == [0.0, 0.5]
    f.close()
    with Common.DataDumper(path) as d:  # read mode
        assert np.array_equal(d.ncfile.variables["v"][:], want)
        assert d.ncfile.title == b"test"


class _HostSim:
    """A host-only simulator with the BaseSimulator surface runSimulation uses:
    the state after time t is (t + p, 2 t - p, sin(t) p, t) for a fixed
    pattern p."""

    def __init__(self, nx, ny, dx, dy, pattern, scale=1.0, context=None, rho=None):
        self.nx, self.ny, self.dx, self.dy = nx, ny, dx, dy
        self.p = pattern
        self.t, self.nt = 0.0, 0
        self.calls = []

    def state(self, t):
        return [np.float32(t) + self.p, np.float32(2 * t) - self.p, np.float32(np.sin(t)) * self.p,
                np.full_like(self.p, t)]

    def simulate(self, t, dt=None):
        self.calls.append((float(t), dt))
        self.t += float(t)
        self.nt += 3

    def download(self, variables=None):
        s = self.state(self.t)
        return [s[i] for i in (range(4) if variables is None else variables)]

    def check(self):
        pass

    def simTime(self):
        return self.t

    def simSteps(self):
        return self.nt

    def getExtent(self):
        return [0.0, self.nx * self.dx, 0.0, self.ny * self.dy]


def test_run_simulation_writes_the_reference_layout(tmp_path):
    nx, ny = 6, 4
    pattern = np.arange(nx * ny, dtype=np.float32).reshape(ny, nx) / 7.0
    sims = []

    def gen(**kw):
        sims.append(_HostSim(**kw))
        return sims[-1]

    args = {"nx": nx, "ny": ny, "dx": 0.5, "dy": 0.25, "pattern": pattern, "scale": 2.0, "rho": pattern}
    times = np.linspace(0.0, 1.0, 4)
    out, prof, mpi = Common.runSimulation(gen, args, str(tmp_path / "run.nc"), times,
                                          ["rho", None, "rho_v", "E"], dt=0.01)
    assert mpi is None and set(prof["start"]) == {"t_sim_init", "t_nc_write", "t_full_step"}
    sim = sims[0]
    # one simulate() per nonzero interval, with the given dt
    assert [c[1] for c in sim.calls] == [0.01] * 3
    assert np.allclose([c[0] for c in sim.calls], np.diff(times))
    f = netcdf_file(out, "r", mmap=False)
    assert f.dimensions == {"time": 4, "x": nx, "y": ny}
    assert np.array_equal(f.variables["time"][:], times.astype(np.float32))
    assert np.array_equal(f.variables["x"][:], np.linspace(0.0, 3.0, nx).astype(np.float32))
    assert np.array_equal(f.variables["y"][:], np.linspace(0.0, 1.0, ny).astype(np.float32))
    assert set(f.variables) == {"time", "x", "y", "rho", "rho_v", "E"}
    t_acc = 0.0
    for k, t_step in enumerate(np.concatenate([[times[0]], np.diff(times)])):
        t_acc += float(t_step) if t_step > 0 else 0.0
        want = sim.state(t_acc)
        for name, idx in (("rho", 0), ("rho_v", 2), ("E", 3)):
            assert f.variables[name].dimensions == ("time", "y", "x")
            assert np.array_equal(f.variables[name][k], Common.quantize(want[idx], 3)), (name, k)
    meta = json.loads(f.sim_args.decode())
    assert meta["nx"] == nx and meta["scale"] == 2.0 and "rho" not in meta
    assert meta["pattern"][0][:2] == [0.0, pytest.approx(1 / 7.0)]  # arrays other than fields as lists
    assert f.simulator and f.created
    f.close()
    # a second run never overwrites the first
    out2, _, _ = Common.runSimulation(gen, args, str(tmp_path / "run.nc"), times, ["E"])
    assert out2.endswith("run_0000.nc")


def test_run_simulation_requires_save_times(tmp_path):
    with pytest.raises(AssertionError):
        Common.runSimulation(lambda **kw: None, {"nx": 2, "ny": 2}, str(tmp_path / "x.nc"), [])


def test_to_json_stringifies_objects():
    class Ctx:
        def __str__(self):
            return "ctx"
    d = json.loads(Common.toJson({"a": np.arange(3), "b": 1.5, "c": Ctx()}))
    assert d == {"a": [0, 1, 2], "b": 1.5, "c": "ctx"}
