"""The single-grid throughput workloads at full size and their parity crop.

* 512x512 version of the 4096x4096 dam break with 16 Gaussian bumps (same
  generator, radius and bump widths scaled by 1/8, the reference's walls),
  500 steps, each scheme in both arithmetic policies against the float32
  oracle: exact arithmetic bitwise; fast arithmetic within 1e-5 relative L1
  per field.  WAF runs exact by default; its fast arithmetic (opt-in) is held
  to max(1e-5, 2 x the float32 oracle's distance from float64), since its
  limiter's upwind selections and ratios turn rounding differences into O(1)
  changes of the limiter weights.  KP07_dimsplit and HLL2 are compared with
  the oracle run in the test; WAF's oracle (minutes at this size) with the
  committed fixture tests/golden/crop512_waf.npz (tests/golden/make_crop_fixture.py),
  whose initial CFL pins the generator the fixture was made from.
* 4096x4096, 500 steps (fast): what holds at any size -- h stays positive and
  finite, and with reflective walls the total mass (sum of h) is conserved to
  rounding (relative 1e-5, float64 sums).
* EE2D_KP07_dimsplit 8192x8192 Rayleigh-Taylor through the one-rank
  MPISimulator, 300 steps: density and pressure stay positive and finite and
  the total mass is conserved (reflective north/south, periodic east/west).
"""

import functools
import os

import numpy as np
import pytest

from oracle import fv_oracle as O
from tests._cases import make_gpu_sim

pytestmark = pytest.mark.gpu

SCHEMES = ("kp07_dimsplit", "hll2", "waf")
R = O.REFLECTIVE
WALLS = O.pack_bc(R, R, R, R)
CROP_STEPS = 500
FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "crop512_waf.npz")


@functools.lru_cache(maxsize=None)
def crop_reference(scheme):
    """(initial fields, dt, float32 oracle fields after 500 steps, float32 vs
    float64 oracle distance per field or None)."""
    from finitevolumegpu_amd.initial_conditions import dam_break_bumps
    h, hu, hv = dam_break_bumps(512, seed=0)
    ora = O.OracleSimulator(scheme, [h, hu, hv], 512, 512, 1.0, 1.0, 9.81, dtype=np.float32)
    dt = np.float32(0.45 * float(ora.cfl))
    if scheme == "waf":
        with np.load(FIXTURE, allow_pickle=False) as z:
            # dt = 0.45 x the initial CFL pins the initial data the fixture was made from
            assert int(z["steps"]) == CROP_STEPS and np.float32(z["dt"]) == dt
            return (h, hu, hv), dt, [z["h"], z["hu"], z["hv"]], z["f32_vs_f64"].copy(), np.float32(z["cfl"])
    for _ in range(CROP_STEPS):
        ora.step(dt)
    return (h, hu, hv), dt, ora.download(), None, np.float32(ora.cfl)


@pytest.mark.parametrize("math", ["exact", "fast"])
@pytest.mark.parametrize("scheme", SCHEMES)
def test_crop_512_against_oracle(ctx, scheme, math):
    print("case crop 512", scheme, math, flush=True)
    fields, dt, ref, d64, cfl_end = crop_reference(scheme)
    sim = make_gpu_sim(ctx, scheme, list(fields), 512, 512, 1.0, 1.0, WALLS, math=math)
    for _ in range(CROP_STEPS):
        sim.step(dt)
    got = sim.download()
    assert got[0].min() > 0.0
    if math == "exact":  # the CFL minimum after the last step, bitwise
        assert np.float32(sim.computeDt() * 2.0) == cfl_end
    for k in range(3):
        if math == "exact":
            assert np.array_equal(got[k], ref[k]), (k, float(np.abs(got[k] - ref[k]).max()))
        else:
            err = O.rel_l1(got[k], ref[k])
            limit = 1e-5 if d64 is None else max(1e-5, 2.0 * float(d64[k]))
            print("field", k, "fast-vs-f32", err, "limit", limit, flush=True)
            assert err <= limit, (k, err, limit)


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
