"""The round-weighted launch plan on the GPU: on domains where the planner
picks it (checked through fvmi_substep_plan), every scheme's fields equal the
equal-segment plan's bitwise, in both arithmetic policies."""

import numpy as np
import pytest

from oracle import fv_oracle as O
from tests._cases import make_gpu_sim, state
from tests.test_plan import plan

pytestmark = pytest.mark.gpu

R, P = O.REFLECTIVE, O.PERIODIC

# sizes where the round-weighted plan applies (its shortest part keeps 16
# rows): 124-column strips for fast Euler, 60 for the others
CASES = [
    ("kp07_dimsplit", "kp07", 4096, 4096, O.pack_bc(R, R, R, R), "fast"),
    ("hll2", "hll2", 4096, 4096, O.pack_bc(P, P, R, R), "fast"),
    ("ee2d_kp07_dimsplit", "ee2d_exact", 2400, 3000, O.pack_bc(R, R, P, P), "exact"),
    ("ee2d_kp07_dimsplit", "ee2d", 3001, 4032, O.pack_bc(R, R, P, P), "fast"),
    ("kp07_dimsplit", "kp07_exact", 3001, 3600, O.pack_bc(R, P, P, R), "exact"),
]


@pytest.mark.parametrize("scheme,key,nx,ny,bc,math", CASES)
def test_round_plan_equals_equal_segments(ctx, scheme, key, nx, ny, bc, math):
    from finitevolumegpu_amd.HipContext import HipContext
    cus = ctx.cu_count
    segs = plan(key, nx, ny, cu_count=cus)
    strips = -(-nx // (124 if key == "ee2d" else 60))
    assert len(segs) == 4 * cus * 4 and (segs[:, 0] >= 0).sum() % 4 == 0, "round plan expected"
    print("case round plan", scheme, nx, ny, bc, math, flush=True)
    fields = state(scheme, nx, ny, seed=3)
    a = make_gpu_sim(ctx, scheme, fields, nx, ny, 1.0, 1.0, bc, math=math)
    b = make_gpu_sim(ctx, scheme, fields, nx, ny, 1.0, 1.0, bc, math=math, tile_rows=64, waves_per_block=4)
    dt = np.float32(0.05) if scheme == "ee2d_kp07_dimsplit" else np.float32(0.9 * a.computeDt())
    for _ in range(3):
        a.step(dt)
        b.step(dt)
    for k, (x, y) in enumerate(zip(a.download(), b.download())):
        assert np.array_equal(x, y), (k, float(np.abs(x - y).max()))
    assert a.computeDt() == b.computeDt()
    assert strips > 0


@pytest.mark.parametrize("key,order,math", [("lxf", 1, "fast"), ("hll", 1, "exact"), ("force", 1, "fast"),
                                             ("kp07u", 2, "fast"), ("kp07u", 2, "exact")])
def test_unsplit_round_plan_equals_equal_segments(ctx, key, order, math):
    from finitevolumegpu_amd.FORCE import FORCE
    from finitevolumegpu_amd.HLL import HLL
    from finitevolumegpu_amd.KP07 import KP07
    from finitevolumegpu_amd.LxF import LxF
    from finitevolumegpu_amd.Simulator import BoundaryCondition
    nx, ny = 4096, 2048
    segs = plan(key, nx, ny, cu_count=ctx.cu_count)
    if key in ("hll", "kp07u"):
        assert len(segs) == 4 * ctx.cu_count * 4, "round plan expected"
    print("case unsplit round plan", key, math, flush=True)
    fields = state("kp07_dimsplit", nx, ny, seed=5)
    bc = BoundaryCondition(BoundaryCondition.getTypes(O.pack_bc(R, P, P, R)))

    def make(**kw):
        if key == "kp07u":
            return KP07(ctx, *fields, nx, ny, 1.0, 1.0, 9.81, order=order, boundary_conditions=bc, math=math, **kw)
        cls = {"lxf": LxF, "hll": HLL, "force": FORCE}[key]
        return cls(ctx, *fields, nx, ny, 1.0, 1.0, 9.81, boundary_conditions=bc, math=math, **kw)

    a = make()
    b = make(tile_rows=64, waves_per_block=4)
    dt = np.float32(0.5 * a.computeDt())
    for _ in range(3):
        a.step(dt)
        b.step(dt)
    for k, (x, y) in enumerate(zip(a.download(), b.download())):
        assert np.array_equal(x, y), (k, float(np.abs(x - y).max()))
    assert a.computeDt() == b.computeDt()
