"""Literal block/thread emulation of the reference tile kernels (slow; tiny grids).

TEST INFRASTRUCTURE ONLY (same import rule as ``oracle/fv_oracle.py``).

Where ``fv_oracle.substep`` restates a substep globally, this module follows
the structure of the CUDA kernels block by block:

* ``readBlock`` (GPUSimulators/cuda/common.h:320-351): periodic index wrap
  (handlePeriodicBoundaryX/Y, :247-288) and the clamp ``min(l, ny+3)``;
* the reflective mirroring thread by thread, in the order N, S, E, W
  (common.h:133-239, 291-315), with the row-major shared-memory address the
  CUDA code computes, so a write past the end of a tile row lands in the next
  row exactly as it would in shared memory;
* the two sweeps over the tile (the same sweep routines as the global oracle:
  their index ranges equal the tile loop bounds of common.h:443-475 and
  SWE2D_KP07_dimsplit.cu:38-39, 77-78);
* ``writeBlock`` of the in-domain interior cells (common.h:359-393);
* ``writeCfl``'s per-block minimum (SWECommon.h:481-532).

Agreement between this emulation and the global restatement on grids whose
sizes are not multiples of the block size pins the global form's ghost fill,
sweep extents and write mask.  For ``nx % bw == 1`` or ``ny % bh == 1`` the
reference tile code misplaces one mirrored ghost column/row; the emulation
reproduces that, which is how the test suite documents it.
"""

import numpy as np

from . import fv_oracle as O


def _read_block(mem, nx, ny, bc, bx, by, w, h):
    """readBlock for the full-domain launch (x0=y0=0, x1=nx, y1=ny)."""
    Q = np.empty((h + 4, w + 4), dtype=mem.dtype)
    for j in range(h + 4):
        l = by + j
        if l < O.GC and O.bc_south(bc) == O.PERIODIC:
            l += ny
        elif l >= ny + O.GC and O.bc_north(bc) == O.PERIODIC:
            l -= ny
        l = min(l, ny + 3)
        for i in range(w + 4):
            k = bx + i
            if k < O.GC and O.bc_west(bc) == O.PERIODIC:
                k += nx
            elif k >= nx + O.GC and O.bc_east(bc) == O.PERIODIC:
                k -= nx
            k = min(k, nx + 3)
            Q[j, i] = mem[l, k]
    return Q


def _reflect(Q, nx, ny, bc, bix, biy, w, h, sx, sy):
    """Thread-by-thread reflective mirroring with flat shared-memory addressing."""
    W = w + 4
    flat = Q.reshape(-1)
    d = Q.dtype.type

    def put(j, i, val):
        idx = j * W + i
        if not (0 <= idx < flat.size):
            raise IndexError("shared-memory write outside the tile array")
        flat[idx] = val

    def get(j, i):
        return flat[j * W + i]

    if O.bc_north(bc) == O.REFLECTIVE:
        writes = []
        for ty in range(h):
            for tx in range(w):
                j = ty + 2
                tj = h * biy + j
                for i in range(tx, w + 4, w):
                    if tj == ny + 1:
                        writes.append((j + 1, i, d(sy) * get(j, i)))
                    if tj == ny:
                        writes.append((j + 3, i, d(sy) * get(j, i)))
        for j, i, v in writes:
            put(j, i, v)
    if O.bc_south(bc) == O.REFLECTIVE:
        writes = []
        for ty in range(h):
            for tx in range(w):
                j = ty + 2
                tj = h * biy + j
                for i in range(tx, w + 4, w):
                    if tj == 2:
                        writes.append((j - 1, i, d(sy) * get(j, i)))
                    if tj == 3:
                        writes.append((j - 3, i, d(sy) * get(j, i)))
        for j, i, v in writes:
            put(j, i, v)
    if O.bc_east(bc) == O.REFLECTIVE:
        writes = []
        for ty in range(h):
            for tx in range(w):
                i = tx + 2
                ti = w * bix + i
                for j in range(ty, h + 4, h):
                    if ti == nx + 1:
                        writes.append((j, i + 1, d(sx) * get(j, i)))
                    if ti == nx:
                        writes.append((j, i + 3, d(sx) * get(j, i)))
        for j, i, v in writes:
            put(j, i, v)
    if O.bc_west(bc) == O.REFLECTIVE:
        writes = []
        for ty in range(h):
            for tx in range(w):
                i = tx + 2
                ti = w * bix + i
                for j in range(ty, h + 4, h):
                    if ti == 2:
                        writes.append((j, i - 1, d(sx) * get(j, i)))
                    if ti == 3:
                        writes.append((j, i - 3, d(sx) * get(j, i)))
        for j, i, v in writes:
            put(j, i, v)


def substep(scheme, U0, U1, params, dt_, step, bc, block_width, block_height):
    """One substep, block by block.  Returns the per-block CFL array
    (``cfl_[gridDim.x*blockIdx.y + blockIdx.x]``) reshaped (gy, gx)."""
    ny, nx = U0[0].shape[0] - 4, U0[0].shape[1] - 4
    d = U0[0].dtype.type
    dt_ = d(dt_)
    w, h = block_width, block_height
    gx = -(-nx // w)
    gy = -(-ny // h)
    signs = O.EULER_SIGNS if scheme == "ee2d_kp07_dimsplit" else O.SWE_SIGNS
    cfl = np.empty((gy, gx), dtype=U0[0].dtype)
    for biy in range(gy):
        for bix in range(gx):
            Q = []
            for mem, (sx, sy) in zip(U0, signs):
                q = _read_block(mem, nx, ny, bc, bix * w, biy * h, w, h)
                _reflect(q, nx, ny, bc, bix, biy, w, h, sx, sy)
                Q.append(q)
            if step == 0:
                Q = O._x_sweep(scheme, Q, params, dt_)
                Q = O._y_sweep(scheme, Q, params, dt_)
            else:
                Q = O._y_sweep(scheme, Q, params, dt_)
                Q = O._x_sweep(scheme, Q, params, dt_)
            if scheme == "ee2d_kp07_dimsplit" and params["g"] > d(0.0):
                I = (slice(2, h + 2), slice(2, w + 2))
                rho_v = Q[2][I].copy()
                Q[2][I] = Q[2][I] - params["g"] * Q[0][I] * dt_
                Q[3][I] = Q[3][I] - params["g"] * rho_v * dt_
            # writeBlock: interior cells inside the domain
            j1 = min(h, ny - biy * h)
            i1 = min(w, nx - bix * w)
            for k in range(len(Q)):
                U1[k][2 + biy * h:2 + biy * h + j1, 2 + bix * w:2 + bix * w + i1] = Q[k][2:2 + j1, 2:2 + i1]
            q = [a[2:2 + j1, 2:2 + i1] for a in Q]
            if scheme == "ee2d_kp07_dimsplit":
                c = O.euler_cfl(q[0], q[1], q[2], params["dx"], params["dy"], params["gamma"])
            else:
                c = O.swe_cfl(q[0], q[1], q[2], params["dx"], params["dy"], params["g"])
            cfl[biy, bix] = np.min(c)
    return cfl
