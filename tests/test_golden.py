"""The oracle reproduces the committed fixtures (tests/golden/oracle_small.npz)
bitwise in float32 and to 1e-13 in float64; fixture FP32-vs-FP64 differences
stay at float32 rounding level."""

import os

import numpy as np
import pytest

from oracle import fv_oracle as O
from tests._cases import SCHEMES

HERE = os.path.dirname(os.path.abspath(__file__))
BCS = {"reflective": O.pack_bc(3, 3, 3, 3), "periodic": O.pack_bc(2, 2, 2, 2),
       "mixed": O.pack_bc(3, 2, 0, 3), "outflow": O.pack_bc(4, 4, 4, 4)}


@pytest.fixture(scope="module")
def golden():
    with np.load(os.path.join(HERE, "golden", "oracle_small.npz"), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def _run(scheme, bcname, padded, dt, dtype):
    nv, ny4, nx4 = padded.shape
    nx, ny = nx4 - 4, ny4 - 4
    dx = 1.0 / 16 if scheme == "ee2d_kp07_dimsplit" else 1.0
    g = 0.1 if scheme == "ee2d_kp07_dimsplit" else 9.81
    sim = O.OracleSimulator(scheme, [a[2:-2, 2:-2] for a in padded], nx, ny, dx, dx, g, gamma=1.4,
                            bc=BCS[bcname], dtype=dtype)
    for k in range(nv):
        sim.u0[k][:] = padded[k]
        sim.u1[k][:] = padded[k]
    for _ in range(6):
        sim.step(dt)
    return np.stack(sim.download()), sim.cfl


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("bcname", sorted(BCS))
def test_oracle_matches_fixture(golden, scheme, bcname):
    key = "%s__%s" % (scheme, bcname)
    padded = golden[key + "__input"]
    dt = golden[key + "__dt"]
    out32, cfl = _run(scheme, bcname, padded, dt, np.float32)
    assert np.array_equal(out32, golden[key + "__f32"])
    assert cfl == golden[key + "__cfl"]
    out64, _ = _run(scheme, bcname, padded, dt, np.float64)
    assert np.allclose(out64, golden[key + "__f64"], rtol=1e-13, atol=1e-13)
    for k in range(out32.shape[0]):
        ref = golden[key + "__f64"][k]
        scale = np.abs(ref).mean() + 1e-3
        assert np.abs(out32[k] - ref).mean() / scale < 1e-4
