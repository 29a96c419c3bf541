"""GPU parity of the reference's other shallow-water schemes (LxF, HLL, FORCE
and the unsplit KP07 with RK2) against oracle/fv_oracle_unsplit.py on the same
seeded inputs, through the public Simulator API:

* math='exact': bitwise equality of every field and of the CFL minimum;
* math='fast' : relative L1 difference <= 1e-5 per field.
"""

import numpy as np
import pytest

from oracle import fv_oracle as O
from oracle import fv_oracle_unsplit as U
from tests._cases import swe_random_state

pytestmark = pytest.mark.gpu

R, P, D, F = O.REFLECTIVE, O.PERIODIC, O.DIRICHLET, O.OUTFLOW
BCS = {
    "reflective": O.pack_bc(R, R, R, R),
    "periodic": O.pack_bc(P, P, P, P),
    "mixed": O.pack_bc(R, P, D, R),
    "mixed2": O.pack_bc(P, R, R, P),
    "outflow": O.pack_bc(F, F, F, F),
}
SCHEMES = ["lxf", "hll", "force", "kp07"]
FAST_TOL = 1e-5


def _gpu(ctx, scheme, fields, nx, ny, bc, math, order=2, tile_rows=16, waves_per_block=2, xcd_remap=True):
    from finitevolumegpu_amd.FORCE import FORCE
    from finitevolumegpu_amd.HLL import HLL
    from finitevolumegpu_amd.KP07 import KP07
    from finitevolumegpu_amd.LxF import LxF
    from finitevolumegpu_amd.Simulator import BoundaryCondition
    b = BoundaryCondition(BoundaryCondition.getTypes(bc))
    kw = dict(boundary_conditions=b, math=math, tile_rows=tile_rows, waves_per_block=waves_per_block,
              xcd_remap=xcd_remap)
    if scheme == "kp07":
        return KP07(ctx, *fields, nx, ny, 1.0, 1.0, 9.81, order=order, **kw)
    cls = {"lxf": LxF, "hll": HLL, "force": FORCE}[scheme]
    return cls(ctx, *fields, nx, ny, 1.0, 1.0, 9.81, **kw)


def _frame(fields, gc):
    """Padded arrays whose ghost frame holds a valid state (Dirichlet edges read
    ghost memory), written into both buffers of both simulators."""
    out = []
    for k, f in enumerate(fields):
        a = np.full((f.shape[0] + 2 * gc, f.shape[1] + 2 * gc), [1.05, 0.01, -0.01][k], dtype=np.float32)
        a[gc:-gc, gc:-gc] = f
        out.append(a)
    return out


def _pair(ctx, scheme, nx, ny, bc, steps, math, order=2, seed=3, **kw):
    fields = swe_random_state(nx, ny, seed)
    sim = _gpu(ctx, scheme, fields, nx, ny, bc, math, order=order, **kw)
    ora = U.OracleUnsplitSimulator(scheme, fields, nx, ny, 1.0, 1.0, 9.81, order=order, bc=bc, dtype=np.float32)
    gc = U.SCHEME_GC[scheme]
    padded = _frame(fields, gc)
    for k, a in enumerate(padded):
        for buf in (sim.u0, sim.u1):
            buf[k].upload(sim.stream, a, extent=[0, 0, nx + 2 * gc, ny + 2 * gc])
        ora.u0[k][:] = a
        ora.u1[k][:] = a
    sim.synchronize()
    assert np.float32(sim.computeDt() / sim.dt_factor) == ora.cfl
    dt = np.float32(0.3 * float(ora.cfl))
    for _ in range(steps):
        sim.step(dt)
        ora.step(dt)
    return sim, ora


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("bcname", sorted(BCS))
def test_exact_bitwise(ctx, scheme, bcname):
    print("case", scheme, bcname, flush=True)
    nx, ny = 131, 77  # partial strips (62 / 60 columns) and segments
    sim, ora = _pair(ctx, scheme, nx, ny, BCS[bcname], 5, "exact")
    for k, (a, b) in enumerate(zip(sim.download(), ora.download())):
        assert np.array_equal(a, b), "%s %s field %d: max |diff| %g" % (scheme, bcname, k, np.abs(a - b).max())
    assert np.float32(sim.computeDt() / sim.dt_factor) == ora.cfl


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("bcname", ["reflective", "mixed"])
def test_fast_tolerance(ctx, scheme, bcname):
    print("case fast", scheme, bcname, flush=True)
    sim, ora = _pair(ctx, scheme, 131, 77, BCS[bcname], 20, "fast")
    for k, (a, b) in enumerate(zip(sim.download(), ora.download())):
        assert O.rel_l1(a, b) <= FAST_TOL, (scheme, bcname, k, O.rel_l1(a, b))


@pytest.mark.parametrize("math", ["exact", "fast"])
def test_kp07_first_order(ctx, math):
    print("case kp07 order 1", math, flush=True)
    sim, ora = _pair(ctx, "kp07", 70, 65, BCS["mixed2"], 6, math, order=1)
    for k, (a, b) in enumerate(zip(sim.download(), ora.download())):
        if math == "exact":
            assert np.array_equal(a, b), (k, float(np.abs(a - b).max()))
        else:
            assert O.rel_l1(a, b) <= FAST_TOL
    assert sim.dt_factor == 1.0


@pytest.mark.parametrize("scheme", SCHEMES)
def test_launch_shapes_agree(ctx, scheme):
    """Every launch shape gives bitwise the same state (both policies)."""
    nx, ny = 150, 97
    fields = swe_random_state(nx, ny, 11)
    for math in ("exact", "fast"):
        ref = _gpu(ctx, scheme, fields, nx, ny, BCS["mixed2"], math, tile_rows=16, waves_per_block=4)
        outs = []
        for tr, wpb, xcd in ((1, 1, True), (7, 3, False), (64, 4, True), (200, 2, True), (0, 4, True)):
            sim = _gpu(ctx, scheme, fields, nx, ny, BCS["mixed2"], math, tile_rows=tr, waves_per_block=wpb,
                       xcd_remap=xcd)
            outs.append(sim)
        dt = np.float32(0.05)
        for _ in range(3):
            ref.step(dt)
            for s in outs:
                s.step(dt)
        want = ref.download()
        for s in outs:
            for k, (a, b) in enumerate(zip(s.download(), want)):
                assert np.array_equal(a, b), (scheme, math, k)


@pytest.mark.parametrize("scheme", SCHEMES)
def test_simulate_dt_policy(ctx, scheme):
    """simulate() through the Simulator API (dt refreshed every 100 steps from
    the device CFL, last step shortened) equals the oracle stepped with the
    same dt sequence, bitwise in exact arithmetic."""
    nx, ny = 64, 48
    fields = swe_random_state(nx, ny, 5)
    bc = BCS["reflective"]
    sim = _gpu(ctx, scheme, fields, nx, ny, bc, "exact")
    dts = []
    orig = sim.step

    def step(dt):
        dts.append(np.float32(dt))
        orig(dt)
    sim.step = step
    t_end = 130 * 0.9 * float(sim.computeDt())  # about 130 steps: one refresh past the first
    sim.simulate(t_end)
    assert len(dts) > 100
    ora = U.OracleUnsplitSimulator(scheme, fields, nx, ny, 1.0, 1.0, 9.81, bc=bc, dtype=np.float32)
    for i, dt in enumerate(dts):
        if i % 100 == 0 and i + 1 < len(dts):
            want = np.float32(float(ora.cfl) * {"lxf": 0.5, "hll": 0.5, "force": 1.0, "kp07": 0.5}[scheme] * 0.9)
            assert dt == want, (i, dt, want)
        ora.step(dt)
    for k, (a, b) in enumerate(zip(sim.download(), ora.download())):
        assert np.array_equal(a, b), (scheme, k)


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("shape", [(2, 2), (3, 5), (60, 2), (62, 3), (63, 4), (125, 2)])
def test_exact_bitwise_edge_shapes(ctx, scheme, shape):
    """Smallest grids and widths at the 62- and 60-column strip boundaries."""
    nx, ny = shape
    print("case edge", scheme, shape, flush=True)
    for bcname in ("reflective", "mixed"):
        sim, ora = _pair(ctx, scheme, nx, ny, BCS[bcname], 3, "exact", tile_rows=1)
        for k, (a, b) in enumerate(zip(sim.download(), ora.download())):
            assert np.array_equal(a, b), (scheme, shape, bcname, k, float(np.abs(a - b).max()))


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("shape", [(61, 9), (62, 5), (63, 7), (125, 11)])
def test_ghost_frames_never_written(ctx, scheme, shape):
    """Only interior cells are written; the halo lanes' stores (and the second
    RK2 stage's reads) use an out-of-range offset.  The ghost frame of both
    buffers keeps its contents, at widths that end a strip partly filled."""
    nx, ny = shape
    print("case ghost frame", scheme, shape, flush=True)
    gc = U.SCHEME_GC[scheme]
    fields = swe_random_state(nx, ny, 9)
    padded = _frame(fields, gc)
    for bcname in ("reflective", "mixed"):
        sim = _gpu(ctx, scheme, fields, nx, ny, BCS[bcname], "fast", tile_rows=4)
        for k, a in enumerate(padded):
            for buf in (sim.u0, sim.u1):
                buf[k].upload(sim.stream, a, extent=[0, 0, nx + 2 * gc, ny + 2 * gc])
        sim.synchronize()
        dt = np.float32(0.3 * sim.computeDt() / sim.dt_factor)
        for _ in range(3):
            sim.step(dt)
        sim.synchronize()
        frame = np.ones((ny + 2 * gc, nx + 2 * gc), dtype=bool)
        frame[gc:-gc, gc:-gc] = False
        for buf in (sim.u0, sim.u1):
            for k in range(3):
                full = buf[k].download(sim.stream, extent=[0, 0, nx + 2 * gc, ny + 2 * gc])
                assert np.array_equal(full[frame], padded[k][frame]), (scheme, shape, bcname, k)
