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

CASES = [
    ("kp07_dimsplit", "kp07", 4096, 1536, O.pack_bc(R, R, R, R), "fast"),
    ("hll2", "hll2", 4096, 1536, O.pack_bc(P, P, R, R), "fast"),
    ("ee2d_kp07_dimsplit", "ee2d", 2400, 2600, O.pack_bc(R, R, P, P), "exact"),
    ("kp07_dimsplit", "kp07", 3001, 2999, O.pack_bc(R, P, P, R), "exact"),
]


@pytest.mark.parametrize("scheme,key,nx,ny,bc,math", CASES)
def test_round_plan_equals_equal_segments(ctx, scheme, key, nx, ny, bc, math):
    from finitevolumegpu_amd.HipContext import HipContext
    cus = ctx.cu_count
    segs = plan(key, nx, ny, cu_count=cus)
    strips = -(-nx // 60)
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
