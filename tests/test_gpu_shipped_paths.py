"""Paths the library ships besides the plain substep loop, on the GPU:

* the Euler sub-area launches of the reference's overlap split
  (EE2D_KP07_dimsplit.py:140-266): the four edge launches followed by the
  interior launch equal a full-domain substep bitwise, CFL included;
* the autotuning hook (Simulator.py:150-154 asks the context's Autotuner,
  Autotuner.py:104-142): a HipContext with autotuning benchmarks the launch
  shapes once, caches them, and the tuned simulator still equals the float32
  oracle bitwise in exact arithmetic;
* failure detection (Common.py:772-779, Simulator.py:210-217): a non-finite
  interior value makes check() raise AssertionError, for two- and one-ghost-
  cell fields (a one-ghost-cell width whose pitch has no room for two), and
  simulate() re-raises it with the step and time appended;
* the fast arithmetic's CFL minimum (computeDt, KP07_dimsplit.py:135-137)
  within 1e-5 relative of the float32 oracle's.
"""

import numpy as np
import pytest

from oracle import fv_oracle as O
from tests._cases import make_gpu_sim, make_oracle, state

pytestmark = pytest.mark.gpu

R, P = O.REFLECTIVE, O.PERIODIC


@pytest.mark.parametrize("math", ["exact", "fast"])
@pytest.mark.parametrize("bc", [O.pack_bc(R, R, P, P), O.pack_bc(R, R, R, R), O.pack_bc(P, P, R, R)])
def test_euler_subarea_split_equals_full_substep(ctx, bc, math):
    nx, ny = 133, 97
    fields = state("ee2d_kp07_dimsplit", nx, ny, 31)
    full = make_gpu_sim(ctx, "ee2d_kp07_dimsplit", fields, nx, ny, 1.0 / 64, 1.0 / 64, bc, math=math)
    split = make_gpu_sim(ctx, "ee2d_kp07_dimsplit", fields, nx, ny, 1.0 / 64, 1.0 / 64, bc, math=math)
    dt = np.float32(0.3 * full.computeDt())
    print("case euler subarea", bc, math, flush=True)
    for _ in range(6):
        full.step(dt)
        for i in range(2):
            split.substep(dt, i, external=True, internal=False)
            split.substep(dt, i, external=False, internal=True)
            split.swapBuffers()
        split.t += float(dt)
        split.nt += 1
    for k, (a, b) in enumerate(zip(split.download(), full.download())):
        assert np.array_equal(a, b), (k, float(np.abs(a - b).max()))
    assert split.computeDt() == full.computeDt()


def test_autotuned_context_benchmarks_once_and_stays_exact(ctx, tmp_path, monkeypatch):
    from finitevolumegpu_amd.Autotuner import Autotuner
    from finitevolumegpu_amd.HipContext import HipContext
    from finitevolumegpu_amd.KP07_dimsplit import KP07_dimsplit
    tuned = HipContext(device=0, autotuning=True)
    assert isinstance(tuned.autotuner, Autotuner)
    tuned.autotuner = Autotuner(nx=256, ny=256, tile_rows=(0, 16, 40), waves_per_block=(2, 4),
                                filename=str(tmp_path / "autotuning_data_test.npz"), timesteps=4)
    calls = {"n": 0}
    orig = Autotuner.run_benchmark

    def counting(*args, **kw):
        calls["n"] += 1
        return orig(*args, **kw)
    monkeypatch.setattr(Autotuner, "run_benchmark", staticmethod(counting))
    nx, ny = 150, 110
    fields = state("kp07_dimsplit", nx, ny, 41)
    print("case autotuned context", flush=True)
    sim = KP07_dimsplit(tuned, *fields, nx, ny, 1.0, 1.0, 9.81, math="exact")
    assert calls["n"] == 6  # every (tile_rows, waves_per_block) pair once
    peak = tuned.autotuner.get_peak_performance(KP07_dimsplit)
    assert peak["tile_rows"] in (0, 16, 40) and peak["waves_per_block"] in (2, 4) and peak["megacells"] > 0
    assert (sim.launch.tile_rows, sim.launch.waves_per_block) == (peak["tile_rows"], peak["waves_per_block"])
    assert (tmp_path / "autotuning_data_test.npz").exists()
    # a second simulator and a fresh tuner on the same cache file do not benchmark again
    KP07_dimsplit(tuned, *fields, nx, ny, 1.0, 1.0, 9.81, math="exact")
    tuned.autotuner = Autotuner(nx=256, ny=256, tile_rows=(0, 16, 40), waves_per_block=(2, 4),
                                filename=str(tmp_path / "autotuning_data_test.npz"))
    assert tuned.autotuner.get_peak_performance(KP07_dimsplit) == peak
    assert calls["n"] == 6
    ora = make_oracle("kp07_dimsplit", fields, nx, ny, 1.0, 1.0, O.pack_bc(R, R, R, R))
    dt = np.float32(0.4 * float(ora.cfl))
    for _ in range(8):
        sim.step(dt)
        ora.step(dt)
    for k, (a, b) in enumerate(zip(sim.download(), ora.download())):
        assert np.array_equal(a, b), k


@pytest.mark.parametrize("scheme,nx", [("kp07_dimsplit", 126), ("lxf", 126), ("lxf", 62), ("hll", 254)])
def test_check_raises_on_a_non_finite_interior_cell(ctx, scheme, nx):
    """A NaN at interior cell (0, 0), the corner a window shifted by one ghost
    cell would miss; the one-ghost-cell widths 126, 62 and 254 have pitches
    of (nx + 2) * 4 rounded to 256 bytes, below (nx + 4) * 4."""
    from finitevolumegpu_amd.HLL import HLL
    from finitevolumegpu_amd.KP07_dimsplit import KP07_dimsplit
    from finitevolumegpu_amd.LxF import LxF
    ny = 40
    h = np.ones((ny, nx), np.float32)
    z = np.zeros((ny, nx), np.float32)
    cls = {"kp07_dimsplit": KP07_dimsplit, "lxf": LxF, "hll": HLL}[scheme]
    sim = cls(ctx, h, z, z.copy(), nx, ny, 1.0, 1.0, 9.81)
    sim.check()  # sane data pass
    field = sim.u0[0]
    bad = h.copy()
    bad[0, 0] = np.nan
    field.upload(sim.stream, bad)
    with pytest.raises(AssertionError, match="NaN"):
        sim.check()
    bad = h.copy()
    bad[ny - 1, nx - 1] = np.inf  # the opposite corner
    field.upload(sim.stream, bad)
    with pytest.raises(AssertionError):
        sim.check()
    # ghost cells are not checked: a NaN in the ghost frame alone passes
    field.upload(sim.stream, h)
    padded = np.ones(field.shape, np.float32)
    padded[0, 0] = np.nan
    field.upload(sim.stream, padded, extent=[0, 0, field.shape[1], field.shape[0]])
    sim.check()


def test_simulate_reports_step_and_time_of_a_blow_up(ctx, monkeypatch):
    from finitevolumegpu_amd import Common
    from finitevolumegpu_amd.KP07_dimsplit import KP07_dimsplit
    n = 64
    h, hu, hv = state("kp07_dimsplit", n, n, 51)
    h = np.array(h)
    h[10, 20] = np.nan
    sim = KP07_dimsplit(ctx, h, hu, hv, n, n, 1.0, 1.0, 9.81)
    # a progress string (and with it check()) after every step
    monkeypatch.setattr(Common.ProgressPrinter, "getPrintString", lambda self, t: "progress")
    with pytest.raises(AssertionError) as err:
        sim.simulate(1.0, dt=0.01)
    text = " ".join(str(a) for a in err.value.args)
    assert "NaN" in text and "Step=1" in text and "time=0.01" in text


@pytest.mark.parametrize("scheme", ["kp07_dimsplit", "hll2", "waf", "ee2d_kp07_dimsplit"])
def test_fast_cfl_minimum_matches_oracle(ctx, scheme):
    """computeDt() of the fast arithmetic (the per-cell rate kept as a maximum,
    one reciprocal per wave) within 1e-5 relative of the float32 oracle's
    CFL minimum, after every step of a short run."""
    nx, ny = 123, 77
    fields = state(scheme, nx, ny, 61)
    dx = 1.0 / 64 if scheme == "ee2d_kp07_dimsplit" else 1.0
    bc = O.pack_bc(R, R, P, P)
    sim = make_gpu_sim(ctx, scheme, fields, nx, ny, dx, dx, bc, math="fast")
    ora = make_oracle(scheme, fields, nx, ny, dx, dx, bc)
    dt = np.float32(0.4 * float(ora.cfl))
    for _ in range(5):
        sim.step(dt)
        ora.step(dt)
        want = float(ora.compute_dt())
        got = float(sim.computeDt())
        assert abs(got - want) <= 1e-5 * want, (got, want)
