"""Writes tests/golden/oracle_small.npz: small seeded cases of every scheme and
boundary mix, with the float32 and float64 oracle results after a few full
steps.  These fixtures pin the oracle against unintended change (the
independent pins are in tests/test_oracle_pins.py) and give the GPU tests a
committed expected output.

    python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import fv_oracle as O  # noqa: E402
from tests._cases import SCHEMES, ghost_frame, state  # noqa: E402

R, P, D, F = O.REFLECTIVE, O.PERIODIC, O.DIRICHLET, O.OUTFLOW
BCS = {"reflective": O.pack_bc(R, R, R, R), "periodic": O.pack_bc(P, P, P, P),
       "mixed": O.pack_bc(R, P, D, R), "outflow": O.pack_bc(F, F, F, F)}
NX, NY, STEPS = 20, 14, 6


def case(scheme, bcname):
    fields = state(scheme, NX, NY, 100 + SCHEMES.index(scheme))
    padded = ghost_frame(fields, len(fields))
    dx = 1.0 / 16 if scheme == "ee2d_kp07_dimsplit" else 1.0
    g = 0.1 if scheme == "ee2d_kp07_dimsplit" else 9.81
    out = {}
    for dtype in (np.float32, np.float64):
        sim = O.OracleSimulator(scheme, fields, NX, NY, dx, dx, g, gamma=1.4, bc=BCS[bcname], dtype=dtype)
        for k, a in enumerate(padded):
            sim.u0[k][:] = a
            sim.u1[k][:] = a
        if dtype == np.float32:
            dt = np.float32(0.45 * float(sim.cfl))
        for _ in range(STEPS):
            sim.step(dt)
        out[np.dtype(dtype).name] = np.stack(sim.download())
        if dtype == np.float32:
            out["cfl"] = np.float32(sim.cfl)
    return np.stack(padded), dt, out


def main():
    data = {}
    for scheme in SCHEMES:
        for bcname in BCS:
            key = "%s__%s" % (scheme, bcname)
            padded, dt, out = case(scheme, bcname)
            data[key + "__input"] = padded
            data[key + "__dt"] = np.float32(dt)
            data[key + "__f32"] = out["float32"]
            data[key + "__f64"] = out["float64"]
            data[key + "__cfl"] = out["cfl"]
    path = os.path.join(HERE, "oracle_small.npz")
    np.savez_compressed(path, **data)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
