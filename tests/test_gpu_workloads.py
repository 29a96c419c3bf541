"""GPU parity on the named workloads: the 256x256 circular dam break (200 steps)
against the float32 oracle, bitwise in exact arithmetic and within 1e-5 relative
L1 in fast arithmetic; positivity of h."""

import numpy as np
import pytest

from oracle import fv_oracle as O
from tests._cases import make_gpu_sim

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c0_oracle():
    from finitevolumegpu_amd.initial_conditions import circular_dam_break
    h, hu, hv = circular_dam_break(256, 256, radius=32.0, seed=0)
    ora = O.OracleSimulator("kp07_dimsplit", [h, hu, hv], 256, 256, 1.0, 1.0, 9.81, dtype=np.float32)
    dt = np.float32(0.45 * float(ora.cfl))
    for _ in range(200):
        ora.step(dt)
    return (h, hu, hv), dt, ora.download(), ora.cfl


@pytest.mark.parametrize("math", ["exact", "fast"])
def test_c0_dam_break_200_steps(ctx, c0_oracle, math):
    fields, dt, ref, ref_cfl = c0_oracle
    print("case c0", math, flush=True)
    sim = make_gpu_sim(ctx, "kp07_dimsplit", list(fields), 256, 256, 1.0, 1.0,
                       O.pack_bc(O.REFLECTIVE, O.REFLECTIVE, O.REFLECTIVE, O.REFLECTIVE), math=math)
    for _ in range(200):
        sim.step(dt)
    got = sim.download()
    assert got[0].min() > 0.0
    for k in range(3):
        if math == "exact":
            assert np.array_equal(got[k], ref[k]), (k, float(np.abs(got[k] - ref[k]).max()))
        else:
            assert O.rel_l1(got[k], ref[k]) <= 1e-5, (k, O.rel_l1(got[k], ref[k]))
    if math == "exact":
        assert np.float32(sim.computeDt() * 2.0) == ref_cfl


def _rt_case(n=96, steps=60):
    from finitevolumegpu_amd.initial_conditions import rayleigh_taylor
    rho, ru, rv, E, dx, dy = rayleigh_taylor(n, n, seed=1)
    bc = O.pack_bc(O.REFLECTIVE, O.REFLECTIVE, O.PERIODIC, O.PERIODIC)
    dt = np.float32(0.45 * dx / np.sqrt(1.4 * 2.6))
    return [rho, ru, rv, E], dx, dy, bc, dt, steps


@pytest.fixture(scope="module")
def rt_oracles():
    fields, dx, dy, bc, dt, steps = _rt_case()
    n = fields[0].shape[0]
    out = {}
    for dtype in (np.float32, np.float64):
        ora = O.OracleSimulator("ee2d_kp07_dimsplit", fields, n, n, dx, dy, 0.1, gamma=1.4, bc=bc, dtype=dtype)
        for _ in range(steps):
            ora.step(dt)
        out[np.dtype(dtype).name] = ora.download()
    return out


@pytest.mark.parametrize("math", ["exact", "fast"])
def test_rayleigh_taylor_euler(ctx, rt_oracles, math):
    """Euler with gravity, Rayleigh-Taylor setup (reflective y, periodic x), 60 steps:
    exact = bitwise, fast within 1e-5 relative L1 of the float32 oracle for the
    density and energy; the momenta start at zero, so their error is compared
    with the float32-vs-float64 oracle difference; rho and p stay positive."""
    fields, dx, dy, bc, dt, steps = _rt_case()
    n = fields[0].shape[0]
    print("case rt", math, flush=True)
    sim = make_gpu_sim(ctx, "ee2d_kp07_dimsplit", fields, n, n, dx, dy, bc, math=math, g=0.1)
    for _ in range(steps):
        sim.step(dt)
    got = sim.download()
    ref32, ref64 = rt_oracles["float32"], rt_oracles["float64"]
    rho, ru, rv, E = got
    p = 0.4 * (E - 0.5 * (ru * ru + rv * rv) / rho)
    assert rho.min() > 0.0 and p.min() > 0.0
    for k in range(4):
        if math == "exact":
            assert np.array_equal(got[k], ref32[k]), (k, float(np.abs(got[k] - ref32[k]).max()))
        else:
            e_fast = O.rel_l1(got[k], ref32[k])
            e_fp32 = O.rel_l1(ref32[k], ref64[k])
            print("field", k, "fast-vs-f32", e_fast, "f32-vs-f64", e_fp32, flush=True)
            if k in (0, 3):
                assert e_fast <= 1e-5, (k, e_fast)
            else:
                assert e_fast <= max(1e-5, 2.0 * e_fp32), (k, e_fast, e_fp32)


@pytest.mark.parametrize("math", ["exact", "fast"])
@pytest.mark.parametrize("wrap", [(True, False), (True, True)])
def test_mpisimulator_halo_exchange(ctx, math, wrap):
    """One-rank MPISimulator (periodic edges -> Dirichlet + on-device halo
    exchange after every substep) equals the plain simulator with periodic
    boundaries bitwise, and the float32 oracle's halo path."""
    from finitevolumegpu_amd.MPISimulator import MPISimulator
    from tests._cases import state
    nx, ny = 150, 70
    fields = state("ee2d_kp07_dimsplit", nx, ny, 31)
    P, R = O.PERIODIC, O.REFLECTIVE
    wx, wy = wrap
    bc = O.pack_bc(P if wy else R, P if wy else R, P if wx else R, P if wx else R)
    dx = 1.0 / 64
    print("case mpisim", math, wrap, flush=True)
    plain = make_gpu_sim(ctx, "ee2d_kp07_dimsplit", fields, nx, ny, dx, dx, bc, math=math)
    wrapped = MPISimulator(make_gpu_sim(ctx, "ee2d_kp07_dimsplit", fields, nx, ny, dx, dx, bc, math=math))
    D = O.DIRICHLET
    ora = O.OracleSimulator("ee2d_kp07_dimsplit", fields, nx, ny, dx, dx, 0.1, gamma=1.4,
                            bc=O.pack_bc(D if wy else R, D if wy else R, D if wx else R, D if wx else R),
                            halo=wrap)
    dt = np.float32(0.45 * float(ora.cfl))
    for _ in range(8):
        plain.step(dt)
        wrapped.step(dt)
        ora.step(dt)
    for k, (a, b, c) in enumerate(zip(plain.download(), wrapped.download(), ora.download())):
        assert np.array_equal(a, b), (k, float(np.abs(a - b).max()))
        if math == "exact":
            assert np.array_equal(b, c), (k, float(np.abs(b - c).max()))
        elif k in (0, 3):
            assert O.rel_l1(b, c) <= 1e-5
    assert wrapped.computeDt() == plain.computeDt()
