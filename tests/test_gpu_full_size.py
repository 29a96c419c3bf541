"""The single-grid throughput workloads at full size and their parity crop.

* 512x512 version of the 4096x4096 dam break with 16 Gaussian bumps (same
  generator, radius and bump widths scaled by 1/8, the reference's walls):
  KP07_dimsplit and HLL2 against the float32 oracle after 500 steps, WAF
  after 200 (its oracle, float32 and float64, would not finish 500 within a
  test's time limit), bitwise in exact arithmetic; in fast arithmetic within
  1e-5 relative L1 per field -- for WAF within twice the float32 oracle's own
  distance from the float64 oracle where that is larger (its momenta: after
  60 steps the float32 oracle is already 2.2e-5 from float64, as rounding
  differences of a few ulp grow through the limiter's upwind selections).
  Each scheme's oracle runs once, in the first test that needs it.
* 4096x4096, 500 steps (fast): what holds at any size -- h stays positive and
  finite, and with reflective walls the total mass (sum of h) is conserved to
  rounding (relative 1e-5, float64 sums).
* EE2D_KP07_dimsplit 8192x8192 Rayleigh-Taylor through the one-rank
  MPISimulator, 300 steps: density and pressure stay positive and finite and
  the total mass is conserved (reflective north/south, periodic east/west).
"""

import functools

import numpy as np
import pytest

from oracle import fv_oracle as O
from tests._cases import make_gpu_sim

pytestmark = pytest.mark.gpu

SCHEMES = ("kp07_dimsplit", "hll2", "waf")
R = O.REFLECTIVE
WALLS = O.pack_bc(R, R, R, R)


CROP_STEPS = {"kp07_dimsplit": 500, "hll2": 500, "waf": 200}


@functools.lru_cache(maxsize=None)
def crop_oracle(scheme):
    """(initial fields, dt, float32 oracle fields, per-field tolerance)."""
    from finitevolumegpu_amd.initial_conditions import dam_break_bumps
    h, hu, hv = dam_break_bumps(512, seed=0)
    ora = O.OracleSimulator(scheme, [h, hu, hv], 512, 512, 1.0, 1.0, 9.81, dtype=np.float32)
    dt = np.float32(0.45 * float(ora.cfl))
    for _ in range(CROP_STEPS[scheme]):
        ora.step(dt)
    ref = ora.download()
    tol = [1e-5] * 3
    if scheme == "waf":
        o64 = O.OracleSimulator(scheme, [h, hu, hv], 512, 512, 1.0, 1.0, 9.81, dtype=np.float64)
        for _ in range(CROP_STEPS[scheme]):
            o64.step(dt)
        tol = [max(1e-5, 2.0 * O.rel_l1(a, b)) for a, b in zip(ref, o64.download())]
    return (h, hu, hv), dt, ref, tol


@pytest.mark.parametrize("math", ["exact", "fast"])
@pytest.mark.parametrize("scheme", SCHEMES)
def test_crop_512_against_oracle(ctx, scheme, math):
    print("case crop 512", scheme, math, flush=True)
    fields, dt, ref, tol = crop_oracle(scheme)
    sim = make_gpu_sim(ctx, scheme, list(fields), 512, 512, 1.0, 1.0, WALLS, math=math)
    for _ in range(CROP_STEPS[scheme]):
        sim.step(dt)
    got = sim.download()
    assert got[0].min() > 0.0
    for k in range(3):
        if math == "exact":
            assert np.array_equal(got[k], ref[k]), (k, float(np.abs(got[k] - ref[k]).max()))
        else:
            assert O.rel_l1(got[k], ref[k]) <= tol[k], (k, O.rel_l1(got[k], ref[k]), tol[k])


@pytest.mark.parametrize("scheme", SCHEMES)
def test_full_size_4096_invariants(ctx, scheme):
    from finitevolumegpu_amd.initial_conditions import dam_break_bumps
    print("case full size 4096", scheme, flush=True)
    h, hu, hv = dam_break_bumps(4096, seed=0)
    mass0 = float(np.sum(h, dtype=np.float64))
    sim = make_gpu_sim(ctx, scheme, [h, hu, hv], 4096, 4096, 1.0, 1.0, WALLS, math="fast")
    dt = np.float32(sim.computeDt() * 0.9)
    for _ in range(500):
        sim.step(dt)
    out = sim.download()
    assert all(bool(np.isfinite(a).all()) for a in out)
    assert out[0].min() > 0.0
    mass = float(np.sum(out[0], dtype=np.float64))
    assert abs(mass - mass0) <= 1e-5 * mass0, (mass, mass0)
    assert np.isfinite(sim.computeDt()) and sim.computeDt() > 0.0


def test_full_size_8192_euler_invariants(ctx):
    from finitevolumegpu_amd.EE2D_KP07_dimsplit import EE2D_KP07_dimsplit
    from finitevolumegpu_amd.MPISimulator import MPISimulator
    from finitevolumegpu_amd.Simulator import BoundaryCondition
    from finitevolumegpu_amd.initial_conditions import rayleigh_taylor
    print("case full size 8192 euler", flush=True)
    n = 8192
    rho, ru, rv, E, dx, dy = rayleigh_taylor(n, n, seed=0)
    mass0 = float(np.sum(rho, dtype=np.float64))
    T = BoundaryCondition.Type
    bc = BoundaryCondition({"north": T.Reflective, "south": T.Reflective, "east": T.Periodic, "west": T.Periodic})
    sim = MPISimulator(EE2D_KP07_dimsplit(ctx, rho, ru, rv, E, n, n, dx, dy, 0.1, 1.4, boundary_conditions=bc))
    dt = np.float32(0.45 * dx / np.sqrt(1.4 * 2.6))
    for _ in range(300):
        sim.step(dt)
    r, mu, mv, e = sim.download()
    for a in (r, mu, mv, e):
        assert bool(np.isfinite(a).all())
    assert r.min() > 0.0
    p = 0.4 * (e - 0.5 * (mu * mu + mv * mv) / r)
    assert p.min() > 0.0
    mass = float(np.sum(r, dtype=np.float64))
    assert abs(mass - mass0) <= 1e-5 * mass0, (mass, mass0)
