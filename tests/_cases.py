"""Shared helpers for the parity tests: build the same case on the GPU (through
the public Simulator API) and on the oracle."""

import numpy as np

from oracle import fv_oracle as O

SCHEMES = ("kp07_dimsplit", "hll2", "waf", "ee2d_kp07_dimsplit")

R, P, D, F = O.REFLECTIVE, O.PERIODIC, O.DIRICHLET, O.OUTFLOW


def swe_random_state(nx, ny, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:ny, 0:nx] + 0.5
    r = np.hypot(x - 0.4 * nx, y - 0.55 * ny)
    h = 1.0 + 0.5 * (r < 0.25 * min(nx, ny)) + rng.uniform(-1e-3, 1e-3, (ny, nx))
    hu = 0.05 * rng.standard_normal((ny, nx))
    hv = 0.05 * rng.standard_normal((ny, nx))
    return [np.ascontiguousarray(a, dtype=np.float32) for a in (h, hu, hv)]


def euler_random_state(nx, ny, seed, gamma=1.4):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:ny, 0:nx] + 0.5
    rho = np.where(y > 0.5 * ny + 2.0 * np.cos(2 * np.pi * x / nx), 2.0, 1.0) + 0.01 * rng.random((ny, nx))
    u = 0.02 * rng.standard_normal((ny, nx))
    v = 0.02 * rng.standard_normal((ny, nx))
    p = 2.5 + 0.01 * rng.random((ny, nx))
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return [np.ascontiguousarray(a, dtype=np.float32) for a in (rho, rho * u, rho * v, E)]


def make_gpu_sim(ctx, scheme, fields, nx, ny, dx, dy, bc_code, math="fast", tile_rows=None,
                 waves_per_block=None, g=None, gamma=1.4, theta=None, xcd_remap=True):
    from finitevolumegpu_amd.Simulator import BoundaryCondition
    bc = BoundaryCondition(BoundaryCondition.getTypes(bc_code))
    kw = dict(boundary_conditions=bc, tile_rows=tile_rows, waves_per_block=waves_per_block, math=math,
              xcd_remap=xcd_remap)
    if theta is not None:
        kw["theta"] = theta
    if scheme == "kp07_dimsplit":
        from finitevolumegpu_amd.KP07_dimsplit import KP07_dimsplit
        return KP07_dimsplit(ctx, *fields, nx, ny, dx, dy, 9.81 if g is None else g, **kw)
    if scheme == "hll2":
        from finitevolumegpu_amd.HLL2 import HLL2
        return HLL2(ctx, *fields, nx, ny, dx, dy, 9.81 if g is None else g, **kw)
    if scheme == "waf":
        from finitevolumegpu_amd.WAF import WAF
        kw.pop("theta", None)
        return WAF(ctx, *fields, nx, ny, dx, dy, 9.81 if g is None else g, **kw)
    from finitevolumegpu_amd.EE2D_KP07_dimsplit import EE2D_KP07_dimsplit
    return EE2D_KP07_dimsplit(ctx, *fields, nx, ny, dx, dy, 0.1 if g is None else g, gamma, **kw)


def make_oracle(scheme, fields, nx, ny, dx, dy, bc_code, dtype=np.float32, g=None, gamma=1.4, theta=None):
    if g is None:
        g = 0.1 if scheme == "ee2d_kp07_dimsplit" else 9.81
    return O.OracleSimulator(scheme, fields, nx, ny, dx, dy, g, gamma=gamma, theta=theta, bc=bc_code, dtype=dtype)


def state(scheme, nx, ny, seed):
    return euler_random_state(nx, ny, seed) if scheme == "ee2d_kp07_dimsplit" else swe_random_state(nx, ny, seed)


def ghost_frame(fields, nvars):
    """Padded copies of ``fields`` whose ghost frame holds a valid state, so that
    Dirichlet edges (which read the buffers' ghost memory) stay well posed."""
    frame = [1.05, 0.01, -0.01, 6.5][:nvars]
    out = []
    for k, f in enumerate(fields):
        a = np.full((f.shape[0] + 4, f.shape[1] + 4), frame[k], dtype=np.float32)
        a[2:-2, 2:-2] = f
        out.append(a)
    return out


def install_frame(sim, ora, padded):
    """Writes the same padded arrays (ghosts included) into both buffers of the
    GPU simulator and of the oracle."""
    ny, nx = padded[0].shape
    for k, a in enumerate(padded):
        for buf in (sim.u0, sim.u1):
            buf[k].upload(sim.stream, a, extent=[0, 0, nx, ny])
        ora.u0[k][:] = a
        ora.u1[k][:] = a
