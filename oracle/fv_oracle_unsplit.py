"""CPU restatement (NumPy) of the reference's other shallow-water schemes:
Lax-Friedrichs (``LxF``), FORCE, first-order HLL (``HLL``) and the unsplit
second-order Kurganov-Petrova scheme with TVD RK2 (``KP07``).

TEST INFRASTRUCTURE ONLY, like ``fv_oracle``: imported by ``tests/`` and never
by the product path.

Same global padded-array form as ``fv_oracle`` but with ``gc`` ghost cells a
side: LxF, FORCE and HLL use one (GPUSimulators/LxF.py:88-95, FORCE.py:94-101,
HLL.py:85-92 allocate ``ArakawaA2D(nx, ny, 1, 1)``), KP07 two (KP07.py:91-100).
Operation order follows the CUDA source so that ``dtype=np.float32`` is an
IEEE single-precision restatement.  The reference kernels call ``readBlock``
and ``writeBlock`` without the sub-area arguments their ``common.h`` requires
(the same defect as the dimensionally split SWE kernels); the full-domain
semantics ``x0=y0=0, x1=nx, y1=ny`` are used, as for those.
"""

import numpy as np

from oracle.fv_oracle import (OUTFLOW, PERIODIC, REFLECTIVE, SWE_SIGNS, bc_east, bc_north, bc_south, bc_west,
                              minmod_slope, swe_cfl, swe_cu_flux, swe_F, swe_hll_flux)

SCHEME_GC = {"lxf": 1, "force": 1, "hll": 1, "kp07": 2}


def fill_ghosts_gc(mem, signs, bc, nx, ny, gc):
    """Ghost fill with ``gc`` ghost cells a side (common.h:133-351 for any gc):
    periodic wrap of the read index (handlePeriodicBoundaryX/Y :247-288, the
    read clamped to the padded array), then mirror north, south, east, west
    (bc*Reflective :133-239: ghost k mirrors cell 2*gc-1-k, with the sign of
    the normal momentum flipped).  Outflow ghosts copy the edge cell (an
    extension, as in ``fv_oracle``); Dirichlet ghosts keep their memory value.
    For gc=2 this is ``fv_oracle.fill_ghosts``."""
    n, s, e, w = bc_north(bc), bc_south(bc), bc_east(bc), bc_west(bc)
    rows = np.arange(ny + 2 * gc)
    cols = np.arange(nx + 2 * gc)
    rows = np.where((rows < gc) & (s == PERIODIC), rows + ny,
                    np.where((rows >= ny + gc) & (n == PERIODIC), rows - ny, rows))
    cols = np.where((cols < gc) & (w == PERIODIC), cols + nx,
                    np.where((cols >= nx + gc) & (e == PERIODIC), cols - nx, cols))
    out = []
    for a, (sx, sy) in zip(mem, signs):
        q = a[rows][:, cols].copy()
        d = q.dtype.type
        for k in range(gc):
            if n == REFLECTIVE:
                q[ny + gc + k, :] = d(sy) * q[ny + gc - 1 - k, :]
            elif n == OUTFLOW:
                q[ny + gc + k, :] = q[ny + gc - 1, :]
        for k in range(gc):
            if s == REFLECTIVE:
                q[gc - 1 - k, :] = d(sy) * q[gc + k, :]
            elif s == OUTFLOW:
                q[gc - 1 - k, :] = q[gc, :]
        for k in range(gc):
            if e == REFLECTIVE:
                q[:, nx + gc + k] = d(sx) * q[:, nx + gc - 1 - k]
            elif e == OUTFLOW:
                q[:, nx + gc + k] = q[:, nx + gc - 1]
        for k in range(gc):
            if w == REFLECTIVE:
                q[:, gc - 1 - k] = d(sx) * q[:, gc + k]
            elif w == OUTFLOW:
                q[:, gc - 1 - k] = q[:, gc]
        out.append(q)
    return out


# --------------------------------------------------------------------------
# Centred fluxes, SWECommon.h:404-473 (Toro 2001, pp. 163-164)
# --------------------------------------------------------------------------

def lxf_1d_flux(Ql, Qr, g, dx, dt_):
    """LxF_1D_flux, SWECommon.h:404-410: 0.5 (F_l + F_r) + dx/(2 dt) (Q_l - Q_r)."""
    d = Ql[0].dtype.type
    Fl = swe_F(*Ql, g)
    Fr = swe_F(*Qr, g)
    k = dx / (d(2.0) * dt_)
    return tuple(d(0.5) * (Fl[i] + Fr[i]) + k * (Ql[i] - Qr[i]) for i in range(3))


def lxf_2d_flux(Ql, Qr, g, dx, dt_):
    """LxF_2D_flux, SWECommon.h:421-427: the 2-D numerical diffusion dx/(4 dt)."""
    d = Ql[0].dtype.type
    Fl = swe_F(*Ql, g)
    Fr = swe_F(*Qr, g)
    k = dx / (d(4.0) * dt_)
    return tuple(d(0.5) * (Fl[i] + Fr[i]) + k * (Ql[i] - Qr[i]) for i in range(3))


def lxw2_1d_flux(Ql, Qr, g, dx, dt_):
    """LxW2_1D_flux (Richtmyer), SWECommon.h:443-450."""
    d = Ql[0].dtype.type
    Fl = swe_F(*Ql, g)
    Fr = swe_F(*Qr, g)
    k = dt_ / (d(2.0) * dx)
    Q = tuple(d(0.5) * (Ql[i] + Qr[i]) + k * (Fl[i] - Fr[i]) for i in range(3))
    return swe_F(*Q, g)


def force_1d_flux(Ql, Qr, g, dx, dt_):
    """FORCE_1D_flux, SWECommon.h:466-470: the mean of the LxF and Richtmyer fluxes."""
    d = Ql[0].dtype.type
    a = lxf_1d_flux(Ql, Qr, g, dx, dt_)
    b = lxw2_1d_flux(Ql, Qr, g, dx, dt_)
    return tuple(d(0.5) * (a[i] + b[i]) for i in range(3))


def _swap(Q):
    return (Q[0], Q[2], Q[1])


def _x_interfaces(Q, rows, flux):
    """Flux through every x interface (padded columns c, c+1) of ``rows``."""
    return flux(tuple(q[rows, :-1] for q in Q), tuple(q[rows, 1:] for q in Q))


def _y_interfaces(Q, cols, flux):
    """Flux through every y interface (padded rows r, r+1) of ``cols``; the y
    sweep passes (h, hv, hu) and swaps back (e.g. SWE2D_LxF.cu:63-90)."""
    G = flux(_swap(tuple(q[:-1, cols] for q in Q)), _swap(tuple(q[1:, cols] for q in Q)))
    return _swap(G)


# --------------------------------------------------------------------------
# One launch of each kernel
# --------------------------------------------------------------------------

def lxf_substep(U0, U1, g, dx, dy, dt_, bc):
    """LxFKernel, SWE2D_LxF.cu:93-166: x and y fluxes of the same state, one
    update ``Q += (F_w - F_e) dt/dx + (G_s - G_n) dt/dy`` (gc = 1).  Writes the
    interior of U1; returns the CFL minimum of the result."""
    gc = 1
    ny, nx = U0[0].shape[0] - 2 * gc, U0[0].shape[1] - 2 * gc
    d = U0[0].dtype.type
    dt_ = d(dt_)
    Q = fill_ghosts_gc(U0, SWE_SIGNS, bc, nx, ny, gc)
    R, C = slice(gc, ny + gc), slice(gc, nx + gc)
    F = _x_interfaces(Q, R, lambda a, b: lxf_2d_flux(a, b, g, dx, dt_))   # (ny, nx+1)
    G = _y_interfaces(Q, C, lambda a, b: lxf_2d_flux(a, b, g, dy, dt_))   # (ny+1, nx)
    out = []
    for k in range(3):
        q = Q[k][R, C] + ((F[k][:, :-1] - F[k][:, 1:]) * dt_ / dx + (G[k][:-1, :] - G[k][1:, :]) * dt_ / dy)
        U1[k][R, C] = q
        out.append(q)
    return d(np.min(swe_cfl(out[0], out[1], out[2], dx, dy, g)))


def _split_first_order(U0, U1, dx, dy, dt_, bc, g, fx, fy):
    """HLLKernel / FORCEKernel (SWE2D_HLL.cu:100-160, SWE2D_FORCE.cu:88-140):
    x fluxes on every padded row (common.h:443-456 evolveF updates the
    interior columns of every row), then y fluxes of that state on every
    column and the update of the interior rows (evolveG, :464-475)."""
    gc = 1
    ny, nx = U0[0].shape[0] - 2 * gc, U0[0].shape[1] - 2 * gc
    d = U0[0].dtype.type
    dt_ = d(dt_)
    Q = fill_ghosts_gc(U0, SWE_SIGNS, bc, nx, ny, gc)
    R, C = slice(gc, ny + gc), slice(gc, nx + gc)
    F = _x_interfaces(Q, slice(None), fx)  # (ny+2, nx+1)
    for k in range(3):
        Q[k][:, C] = Q[k][:, C] + (F[k][:, :-1] - F[k][:, 1:]) * dt_ / dx
    G = _y_interfaces(Q, slice(None), fy)  # (ny+1, nx+2)
    for k in range(3):
        Q[k][R, :] = Q[k][R, :] + (G[k][:-1, :] - G[k][1:, :]) * dt_ / dy
        U1[k][R, C] = Q[k][R, C]
    return d(np.min(swe_cfl(Q[0][R, C], Q[1][R, C], Q[2][R, C], dx, dy, g)))


def hll_substep(U0, U1, g, dx, dy, dt_, bc):
    return _split_first_order(U0, U1, dx, dy, dt_, bc, g, lambda a, b: swe_hll_flux(a, b, g),
                              lambda a, b: swe_hll_flux(a, b, g))


def force_substep(U0, U1, g, dx, dy, dt_, bc):
    d = U0[0].dtype.type
    dt_ = d(dt_)
    return _split_first_order(U0, U1, dx, dy, dt_, bc, g, lambda a, b: force_1d_flux(a, b, g, dx, dt_),
                              lambda a, b: force_1d_flux(a, b, g, dy, dt_))


def kp07_substep(U0, U1, g, theta, dx, dy, dt_, bc, step, order):
    """KP07Kernel, SWE2D_KP07.cu:120-230 (gc = 2): minmod slopes along x and
    y of the same state (limiters.h:31-41 via minmodSlopeX/Y :103-137), face
    values Q -+ slope/2 with no predictor, central-upwind fluxes (SWECommon.h:
    228-241), one update ``Q += (F_w - F_e) dt/dx + (G_s - G_n) dt/dy``.  The
    second stage of the RK2 order-2 pair writes ``0.5 (U1_old + Q)``
    (:205-214).  The CFL minimum is that of Q before the average (writeCfl
    reads the tile, :222-225)."""
    gc = 2
    ny, nx = U0[0].shape[0] - 2 * gc, U0[0].shape[1] - 2 * gc
    d = U0[0].dtype.type
    dt_ = d(dt_)
    half = d(0.5)
    Q = fill_ghosts_gc(U0, SWE_SIGNS, bc, nx, ny, gc)
    R, C = slice(gc, ny + gc), slice(gc, nx + gc)
    # x: slopes of columns 1..nx+2 on interior rows; interfaces (c, c+1), c = 1..nx+1
    Sx = [minmod_slope(q[R, :-2], q[R, 1:-1], q[R, 2:], theta) for q in Q]  # columns 1..nx+2
    Qm = tuple(Q[k][R, 1:nx + 2] + half * Sx[k][:, 0:nx + 1] for k in range(3))
    Qp = tuple(Q[k][R, 2:nx + 3] - half * Sx[k][:, 1:nx + 2] for k in range(3))
    F = swe_cu_flux(Qm, Qp, g)  # (ny, nx+1)
    # y: slopes of rows 1..ny+2 on interior columns; interfaces (r, r+1), r = 1..ny+1
    Sy = [minmod_slope(q[:-2, C], q[1:-1, C], q[2:, C], theta) for q in Q]  # rows 1..ny+2
    Qm = tuple(Q[k][1:ny + 2, C] + half * Sy[k][0:ny + 1, :] for k in range(3))
    Qp = tuple(Q[k][2:ny + 3, C] - half * Sy[k][1:ny + 2, :] for k in range(3))
    G = _swap(swe_cu_flux(_swap(Qm), _swap(Qp), g))  # (ny+1, nx)
    res = []
    for k in range(3):
        q = Q[k][R, C] + ((F[k][:, :-1] - F[k][:, 1:]) * dt_ / dx + (G[k][:-1, :] - G[k][1:, :]) * dt_ / dy)
        res.append(q)
    for k in range(3):
        if order == 2 and step == 1:
            U1[k][R, C] = half * (U1[k][R, C] + res[k])
        else:
            U1[k][R, C] = res[k]
    return d(np.min(swe_cfl(res[0], res[1], res[2], dx, dy, g)))


def pad_gc(field, dtype, gc):
    """Zero-initialised padded buffer with ``gc`` ghost cells, interior uploaded."""
    f = np.asarray(field)
    ny, nx = f.shape
    out = np.zeros((ny + 2 * gc, nx + 2 * gc), dtype=dtype)
    out[gc:ny + gc, gc:nx + gc] = f.astype(dtype)
    return out


class OracleUnsplitSimulator:
    """Stateful restatement of LxF.py / FORCE.py / HLL.py / KP07.py plus
    BaseSimulator (Simulator.py:176-228): substeps per step (1, or ``order``
    for KP07), swap after every substep, initial CFL from the host arrays, and
    computeDt = min(cfl) times 0.5 (LxF, HLL: LxF.py:123-125, HLL.py:122-124),
    1 (FORCE.py:127-129) or 0.5**(order-1) (KP07.py:136-138)."""

    def __init__(self, scheme, fields, nx, ny, dx, dy, g, theta=1.3, order=2, cfl_scale=0.9, bc=0,
                 dtype=np.float32):
        if scheme not in SCHEME_GC:
            raise ValueError(scheme)
        d = np.dtype(dtype).type
        self.scheme = scheme
        self.gc = SCHEME_GC[scheme]
        self.nx, self.ny = nx, ny
        self.dx, self.dy, self.g, self.theta = d(dx), d(dy), d(g), d(theta)
        self.order = order if scheme == "kp07" else 1
        self.cfl_scale = cfl_scale
        self.bc = bc
        self.u0 = [pad_gc(f, dtype, self.gc) for f in fields]
        self.u1 = [np.zeros_like(a) for a in self.u0]
        h0, hu0, hv0 = (np.asarray(f, dtype=dtype) for f in fields)
        dt_x = np.min(self.dx / (np.abs(hu0 / h0) + np.sqrt(self.g * h0)))
        dt_y = np.min(self.dy / (np.abs(hv0 / h0) + np.sqrt(self.g * h0)))
        self.cfl = d(min(dt_x, dt_y))
        self.t, self.nt = 0.0, 0

    def substep(self, dt, step):
        s = self.scheme
        if s == "lxf":
            c = lxf_substep(self.u0, self.u1, self.g, self.dx, self.dy, dt, self.bc)
        elif s == "force":
            c = force_substep(self.u0, self.u1, self.g, self.dx, self.dy, dt, self.bc)
        elif s == "hll":
            c = hll_substep(self.u0, self.u1, self.g, self.dx, self.dy, dt, self.bc)
        else:
            c = kp07_substep(self.u0, self.u1, self.g, self.theta, self.dx, self.dy, dt, self.bc, step, self.order)
        self.cfl = c
        self.u0, self.u1 = self.u1, self.u0

    def step(self, dt):
        for i in range(self.order):
            self.substep(dt, i)
        self.t += float(dt)
        self.nt += 1

    def compute_dt(self):
        factor = {"lxf": 0.5, "hll": 0.5, "force": 1.0}.get(self.scheme, 0.5 ** (self.order - 1))
        return self.cfl * factor

    def download(self):
        gc = self.gc
        return [a[gc:self.ny + gc, gc:self.nx + gc].copy() for a in self.u0]
