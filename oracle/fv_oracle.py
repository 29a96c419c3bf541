"""CPU restatement (NumPy) of the FiniteVolumeGPU flux-and-update hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: it may be imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py`` and nowhere else.  The product path (``finitevolumegpu_amd``)
never imports it and has no CPU fallback.

The restatement is written in *global padded-array form*: every field is a
``(ny+4, nx+4)`` array with the interior at ``[2:ny+2, 2:nx+2]``, exactly the
buffer layout of ``Common.CudaArray2D`` (GPUSimulators/Common.py:498-531).
One substep is

1. ghost fill (the block-local ``readBlock`` + reflective mirroring of
   GPUSimulators/cuda/common.h:133-351, restated globally),
2. the first sweep over every line of the padded array (the tile's extended
   first sweep, common.h:443-475 loop bounds), then the second sweep,
3. (Euler) the gravity source,
4. the interior written to the *other* buffer, and the CFL minimum.

Operation order follows the CUDA source expression by expression so that the
FP32 path (``dtype=np.float32``) is an IEEE single-precision restatement; the
FP64 path (``dtype=np.float64``) is the same algorithm in double precision.

Parity pins (tests/test_oracle_*.py): known-answer tests for every limiter and
flux, consistency ``flux(Q,Q) == F(Q)``, lake at rest, conservation, mirror
symmetry, the SWASHES Stoker dam-break fixtures, observed convergence order,
and a literal block/thread emulation of the reference tile kernels
(``oracle/tile_emulator.py``) on small odd grids.
"""

import numpy as np

# Boundary condition codes, GPUSimulators/Simulator.py:44-52 and common.h:109-114
DIRICHLET = 0
NEUMANN = 1
PERIODIC = 2
REFLECTIVE = 3
# Zero-gradient ghost copy.  The reference has no outflow type; this code is an
# extension used by the batched ensemble (documented as such).
OUTFLOW = 4

GC = 2  # ghost cells per side, all four schemes (e.g. SWE2D_KP07_dimsplit.cu:147-148)


def bc_north(bc):  # common.h:116-118
    return (int(bc) >> 24) & 0xF


def bc_south(bc):  # common.h:120-122
    return (int(bc) >> 16) & 0xF


def bc_east(bc):  # common.h:124-126
    return (int(bc) >> 8) & 0xF


def bc_west(bc):  # common.h:128-130
    return int(bc) & 0xF


def pack_bc(north, south, east, west):
    """BoundaryCondition.asCodedInt, GPUSimulators/Simulator.py:78-92."""
    return ((north & 0xF) << 24) | ((south & 0xF) << 16) | ((east & 0xF) << 8) | (west & 0xF)


# --------------------------------------------------------------------------
# Limiters (GPUSimulators/cuda/limiters.h, common.h)
# --------------------------------------------------------------------------

def minmod_slope(left, center, right, theta):
    """limiters.h:31-41 (minmodSlope), operation order kept."""
    dt = center.dtype.type
    backward = (center - left) * theta
    central = (right - left) * dt(0.5)
    forward = (right - center) * theta
    one = dt(1.0)
    sb = np.copysign(one, backward)
    sc = np.copysign(one, central)
    sf = np.copysign(one, forward)
    m = np.fmin(np.fmin(np.abs(backward), np.abs(central)), np.abs(forward))
    return dt(0.25) * sb * (sb + sc) * (sc + sf) * m


def generalized_minmod(r, theta):
    """limiters.h:100-102."""
    dt = np.asarray(r).dtype.type
    return np.fmax(dt(0.0), np.fmin(theta * r, np.fmin((dt(1.0) + r) / dt(2.0), theta)))


def desingularize(x, eps):
    """common.h:81-83."""
    dt = np.asarray(x).dtype.type
    return np.copysign(dt(1.0), x) * np.fmax(np.abs(x), np.fmin(x * x / (dt(2.0) * eps) + dt(0.5) * eps, eps))


def limiter_to_waf_limiter(r, c):
    """SWECommon.h:109-111."""
    dt = np.asarray(r).dtype.type
    return dt(1.0) - (dt(1.0) - np.abs(c)) * r


# --------------------------------------------------------------------------
# Shallow water (GPUSimulators/cuda/SWECommon.h)
# --------------------------------------------------------------------------

def swe_F(h, hu, hv, g):
    """SWECommon.h:29-37.  (h, hu, hv) are (depth, normal, tangential)."""
    dt = h.dtype.type
    Fx = hu
    Fy = hu * hu / h + dt(0.5) * g * h * h
    Fz = hu * hv / h
    return Fx, Fy, Fz


def swe_cu_flux(Qm, Qp, g):
    """CentralUpwindFlux, SWECommon.h:228-241.  Qm is the left state."""
    dt = Qm[0].dtype.type
    Fp = swe_F(*Qp, g)
    up = Qp[1] / Qp[0]
    cp = np.sqrt(g * Qp[0])
    Fm = swe_F(*Qm, g)
    um = Qm[1] / Qm[0]
    cm = np.sqrt(g * Qm[0])
    zero = dt(0.0)
    am = np.fmin(np.fmin(um - cm, up - cp), zero)
    ap = np.fmax(np.fmax(um + cm, up + cp), zero)
    apam = ap * am
    den = ap - am
    return tuple(((ap * Fm[k] - am * Fp[k]) + apam * (Qp[k] - Qm[k])) / den for k in range(3))


def swe_hll_flux(Ql, Qr, g):
    """HLL_flux (SWE), SWECommon.h:274-313."""
    dt = Ql[0].dtype.type
    h_l, h_r = Ql[0], Qr[0]
    u_l = Ql[1] / h_l
    u_r = Qr[1] / h_r
    c_l = np.sqrt(g * h_l)
    c_r = np.sqrt(g * h_r)
    h_dag = dt(0.5) * (h_l + h_r) - dt(0.25) * (u_r - u_l) * (h_l + h_r) / (c_l + c_r)
    with np.errstate(invalid="ignore"):
        q_l_tmp = np.sqrt(dt(0.5) * ((h_dag + h_l) * h_dag / (h_l * h_l)))
        q_r_tmp = np.sqrt(dt(0.5) * ((h_dag + h_r) * h_dag / (h_r * h_r)))
    q_l = np.where(h_dag > h_l, q_l_tmp, dt(1.0))
    q_r = np.where(h_dag > h_r, q_r_tmp, dt(1.0))
    S_l = u_l - c_l * q_l
    S_r = u_r + c_r * q_r
    F_l = swe_F(*Ql, g)
    F_r = swe_F(*Qr, g)
    srsl = S_r * S_l
    den = S_r - S_l
    out = []
    for k in range(3):
        with np.errstate(invalid="ignore", divide="ignore"):
            star = (S_r * F_l[k] - S_l * F_r[k] + srsl * (Qr[k] - Ql[k])) / den
        out.append(np.where(S_l >= dt(0.0), F_l[k], np.where(S_r <= dt(0.0), F_r[k], star)))
    return tuple(out)


def compute_h_star(h_l, h_r, u_l, u_r, c_l, c_r, g):
    """SWECommon.h:114-121."""
    dt = h_l.dtype.type
    h_tmp = dt(0.5) * (c_l + c_r) + dt(0.25) * (u_l - u_r)
    return h_tmp * h_tmp / g


def swe_waf_flux(Q_l2, Q_l1, Q_r1, Q_r2, g, dx, dt_):
    """WAF_1D_flux, SWECommon.h:131-210, for the interface between Q_l1 and Q_r1."""
    dt = Q_l1[0].dtype.type
    h_l, h_r = Q_l1[0], Q_r1[0]
    h_l2, h_r2 = Q_l2[0], Q_r2[0]
    u_l = Q_l1[1] / h_l
    u_r = Q_r1[1] / h_r
    u_l2 = Q_l2[1] / h_l2
    u_r2 = Q_r2[1] / h_r2
    v_l = Q_l1[2] / h_l
    v_r = Q_r1[2] / h_r
    v_l2 = Q_l2[2] / h_l2
    v_r2 = Q_r2[2] / h_r2
    c_l = np.sqrt(g * h_l)
    c_r = np.sqrt(g * h_r)
    c_l2 = np.sqrt(g * h_l2)
    c_r2 = np.sqrt(g * h_r2)
    h_dag_l = compute_h_star(h_l2, h_l, u_l2, u_l, c_l2, c_l, g)
    h_dag = compute_h_star(h_l, h_r, u_l, u_r, c_l, c_r, g)
    h_dag_r = compute_h_star(h_r, h_r2, u_r, u_r2, c_r, c_r2, g)
    half = dt(0.5)
    with np.errstate(invalid="ignore"):
        q_l_tmp = np.sqrt(half * ((h_dag + h_l) * h_dag)) / h_l
        q_r_tmp = np.sqrt(half * ((h_dag + h_r) * h_dag)) / h_r
    q_l = np.where(h_dag > h_l, q_l_tmp, dt(1.0))
    q_r = np.where(h_dag > h_r, q_r_tmp, dt(1.0))
    S_l = u_l - c_l * q_l
    S_r = u_r + c_r * q_r
    S_star = (S_l * h_r * (u_r - S_r) - S_r * h_l * (u_l - S_l)) / (h_r * (u_r - S_r) - h_l * (u_l - S_l))
    s_l = h_l * (S_l - u_l) / (S_l - S_star)
    s_r = h_r * (S_r - u_r) / (S_r - S_star)
    one = dt(1.0)
    Q_star_l = (s_l * one, s_l * S_star, s_l * v_l)
    Q_star_r = (s_r * one, s_r * S_star, s_r * v_r)
    F_1 = swe_F(*Q_l1, g)
    F_4 = swe_F(*Q_r1, g)
    F_2 = tuple(F_1[k] + S_l * (Q_star_l[k] - Q_l1[k]) for k in range(3))
    F_3 = tuple(F_4[k] + S_r * (Q_star_r[k] - Q_r1[k]) for k in range(3))
    c_1 = S_l * dt_ / dx
    c_2 = S_star * dt_ / dx
    c_3 = S_r * dt_ / dx
    eps = dt(1.0e-6)
    zero = dt(0.0)
    r_1 = desingularize(np.where(c_1 > zero, h_dag_l - h_l2, h_dag_r - h_r), eps) / desingularize(h_dag - h_l, eps)
    r_2 = desingularize(np.where(c_2 > zero, v_l - v_l2, v_r2 - v_r), eps) / desingularize(v_r - v_l, eps)
    r_3 = desingularize(np.where(c_3 > zero, h_l - h_dag_l, h_r2 - h_dag_r), eps) / desingularize(h_r - h_dag, eps)
    lim = dt(1.9)
    A_1 = np.copysign(one, c_1) * limiter_to_waf_limiter(generalized_minmod(r_1, lim), c_1)
    A_2 = np.copysign(one, c_2) * limiter_to_waf_limiter(generalized_minmod(r_2, lim), c_2)
    A_3 = np.copysign(one, c_3) * limiter_to_waf_limiter(generalized_minmod(r_3, lim), c_3)
    return tuple(half * (F_1[k] + F_4[k])
                 - half * (A_1 * (F_2[k] - F_1[k]) + A_2 * (F_3[k] - F_2[k]) + A_3 * (F_4[k] - F_3[k]))
                 for k in range(3))


def swe_cfl(h, hu, hv, dx, dy, g):
    """Per-cell admissible dt, writeCfl SWECommon.h:497-504."""
    u = hu / h
    v = hv / h
    max_u = dx / (np.abs(u) + np.sqrt(g * h))
    max_v = dy / (np.abs(v) + np.sqrt(g * h))
    return np.fmin(max_u, max_v)


# --------------------------------------------------------------------------
# Euler (GPUSimulators/cuda/EulerCommon.h)
# --------------------------------------------------------------------------

def euler_pressure(Q, gamma):
    """EulerCommon.h:83-90."""
    dt = Q[0].dtype.type
    rho, rho_u, rho_v, E = Q
    return (gamma - dt(1.0)) * (E - dt(0.5) * (rho_u * rho_u + rho_v * rho_v) / rho)


def euler_F(Q, P):
    """EulerCommon.h:93-109.  The second argument is a pressure."""
    rho, rho_u, rho_v, E = Q
    u = rho_u / rho
    return (rho_u, rho_u * u + P, rho_v * u, u * (E + P))


def euler_hll_flux(Q_l, Q_r, gamma):
    """EulerCommon.h:117-160."""
    dt = Q_l[0].dtype.type
    h_l, h_r = Q_l[0], Q_r[0]
    u_l = Q_l[1] / h_l
    u_r = Q_r[1] / h_r
    P_l = euler_pressure(Q_l, gamma)
    P_r = euler_pressure(Q_r, gamma)
    with np.errstate(invalid="ignore"):
        c_l = np.sqrt(gamma * P_l / Q_l[0])
        c_r = np.sqrt(gamma * P_r / Q_r[0])
    h_dag = dt(0.5) * (h_l + h_r) - dt(0.25) * (u_r - u_l) * (h_l + h_r) / (c_l + c_r)
    with np.errstate(invalid="ignore"):
        q_l_tmp = np.sqrt(dt(0.5) * ((h_dag + h_l) * h_dag / (h_l * h_l)))
        q_r_tmp = np.sqrt(dt(0.5) * ((h_dag + h_r) * h_dag / (h_r * h_r)))
    q_l = np.where(h_dag > h_l, q_l_tmp, dt(1.0))
    q_r = np.where(h_dag > h_r, q_r_tmp, dt(1.0))
    S_l = u_l - c_l * q_l
    S_r = u_r + c_r * q_r
    F_l = euler_F(Q_l, P_l)
    F_r = euler_F(Q_r, P_r)
    srsl = S_r * S_l
    den = S_r - S_l
    out = []
    for k in range(4):
        with np.errstate(invalid="ignore", divide="ignore"):
            star = (S_r * F_l[k] - S_l * F_r[k] + srsl * (Q_r[k] - Q_l[k])) / den
        out.append(np.where(S_l >= dt(0.0), F_l[k], np.where(S_r <= dt(0.0), F_r[k], star)))
    return tuple(out)


def euler_cu_flux(Qm, Qp, gamma):
    """EulerCommon.h:171-187."""
    dt = Qm[0].dtype.type
    Pp = euler_pressure(Qp, gamma)
    Fp = euler_F(Qp, Pp)
    up = Qp[1] / Qp[0]
    with np.errstate(invalid="ignore"):
        cp = np.sqrt(gamma * Pp / Qp[0])
    Pm = euler_pressure(Qm, gamma)
    Fm = euler_F(Qm, Pm)
    um = Qm[1] / Qm[0]
    with np.errstate(invalid="ignore"):
        cm = np.sqrt(gamma * Pm / Qm[0])
    zero = dt(0.0)
    am = np.fmin(np.fmin(um - cm, up - cp), zero)
    ap = np.fmax(np.fmax(um + cm, up + cp), zero)
    apam = ap * am
    den = ap - am
    return tuple(((ap * Fm[k] - am * Fp[k]) + apam * (Qp[k] - Qm[k])) / den for k in range(4))


def euler_cfl(rho, rho_u, rho_v, dx, dy, gamma):
    """Per-cell admissible dt, EulerCommon.h:41-49 (uses sqrt(gamma*rho) as written)."""
    u = rho_u / rho
    v = rho_v / rho
    max_u = dx / (np.abs(u) + np.sqrt(gamma * rho))
    max_v = dy / (np.abs(v) + np.sqrt(gamma * rho))
    return np.fmin(max_u, max_v)


# --------------------------------------------------------------------------
# Ghost fill
# --------------------------------------------------------------------------

# Per-variable reflection signs (sign_x, sign_y): SWE2D_KP07_dimsplit.cu:157-159,
# EE2D_KP07_dimsplit.cu:174-177.
SWE_SIGNS = ((1, 1), (-1, 1), (1, -1))
EULER_SIGNS = ((1, 1), (-1, 1), (1, -1), (1, 1))


def fill_ghosts(mem, signs, bc, nx, ny):
    """Values a kernel sees in the padded tile, restated globally.

    common.h:320-351: read with the periodic index wrap (handlePeriodicBoundaryX/Y,
    :247-288), then mirror in the order north, south, east, west
    (handleReflectiveBoundary :291-315, bc*Reflective :133-239).  Dirichlet
    ghosts keep their memory value.  ``mem`` is a list of padded arrays; new
    arrays are returned.
    """
    n, s, e, w = bc_north(bc), bc_south(bc), bc_east(bc), bc_west(bc)
    rows = np.arange(ny + 4)
    cols = np.arange(nx + 4)
    # periodic wrap of the read index (south checked first, then north; same in x)
    rows = np.where((rows < GC) & (s == PERIODIC), rows + ny, np.where((rows >= ny + GC) & (n == PERIODIC), rows - ny, rows))
    cols = np.where((cols < GC) & (w == PERIODIC), cols + nx, np.where((cols >= nx + GC) & (e == PERIODIC), cols - nx, cols))
    out = []
    for a, (sx, sy) in zip(mem, signs):
        q = a[rows][:, cols].copy()
        dt = q.dtype.type
        if n == REFLECTIVE:
            q[ny + 2, :] = dt(sy) * q[ny + 1, :]
            q[ny + 3, :] = dt(sy) * q[ny, :]
        elif n == OUTFLOW:
            q[ny + 2, :] = q[ny + 1, :]
            q[ny + 3, :] = q[ny + 1, :]
        if s == REFLECTIVE:
            q[1, :] = dt(sy) * q[2, :]
            q[0, :] = dt(sy) * q[3, :]
        elif s == OUTFLOW:
            q[1, :] = q[2, :]
            q[0, :] = q[2, :]
        if e == REFLECTIVE:
            q[:, nx + 2] = dt(sx) * q[:, nx + 1]
            q[:, nx + 3] = dt(sx) * q[:, nx]
        elif e == OUTFLOW:
            q[:, nx + 2] = q[:, nx + 1]
            q[:, nx + 3] = q[:, nx + 1]
        if w == REFLECTIVE:
            q[:, 1] = dt(sx) * q[:, 2]
            q[:, 0] = dt(sx) * q[:, 3]
        elif w == OUTFLOW:
            q[:, 1] = q[:, 2]
            q[:, 0] = q[:, 2]
        out.append(q)
    return out


# --------------------------------------------------------------------------
# One-dimensional sweeps along the last axis
# --------------------------------------------------------------------------

def _sweep_reconstructed(Q, theta, dt_, dx, F, flux):
    """Slope-limited sweep (KP07_dimsplit, HLL2, EE2D): SWE2D_KP07_dimsplit.cu:32-69
    (computeFluxF), limiters.h:49-61 (minmodSlopeX), common.h:443-454 (evolveF).

    Updates columns 2..C-3 of every row of each array in ``Q`` (C = padded width).
    """
    d = Q[0].dtype.type
    nv = len(Q)
    C = Q[0].shape[-1]
    half = d(0.5)
    S = [minmod_slope(q[..., :-2], q[..., 1:-1], q[..., 2:], theta) for q in Q]  # cells 1..C-2
    QL = tuple(Q[k][..., 1:-1] - half * S[k] for k in range(nv))
    QR = tuple(Q[k][..., 1:-1] + half * S[k] for k in range(nv))
    FL = F(QL)
    FR = F(QR)
    kk = dt_ / (d(2.0) * dx)
    # Q_l_bar of the interface (i, i+1) is the right face of cell i; Q_r_bar the left face of cell i+1.
    QbarR = tuple(QR[k] + kk * (FL[k] - FR[k]) for k in range(nv))
    QbarL = tuple(QL[k] + kk * (FL[k] - FR[k]) for k in range(nv))
    Fi = flux(tuple(a[..., :-1] for a in QbarR), tuple(a[..., 1:] for a in QbarL))  # interfaces 1.5 .. C-2.5
    out = []
    for k in range(nv):
        q = Q[k].copy()
        q[..., 2:C - 2] = q[..., 2:C - 2] + (Fi[k][..., :-1] - Fi[k][..., 1:]) * dt_ / dx
        out.append(q)
    return out


def _sweep_waf(Q, g, dt_, dx):
    """WAF sweep: SWE2D_WAF.cu:35-54 (computeFluxF) + common.h:443-454 (evolveF)."""
    C = Q[0].shape[-1]
    Ql2 = tuple(q[..., 0:C - 3] for q in Q)
    Ql1 = tuple(q[..., 1:C - 2] for q in Q)
    Qr1 = tuple(q[..., 2:C - 1] for q in Q)
    Qr2 = tuple(q[..., 3:C] for q in Q)
    Fi = swe_waf_flux(Ql2, Ql1, Qr1, Qr2, g, dx, dt_)
    out = []
    for k in range(3):
        q = Q[k].copy()
        q[..., 2:C - 2] = q[..., 2:C - 2] + (Fi[k][..., :-1] - Fi[k][..., 1:]) * dt_ / dx
        out.append(q)
    return out


def _x_sweep(scheme, Q, p, dt_):
    d = Q[0].dtype.type
    if scheme == "waf":
        return _sweep_waf(Q, p["g"], dt_, p["dx"])
    if scheme in ("kp07_dimsplit", "hll2"):
        g = p["g"]
        F = lambda q: swe_F(*q, g)
        flux = (lambda a, b: swe_cu_flux(a, b, g)) if scheme == "kp07_dimsplit" else (lambda a, b: swe_hll_flux(a, b, g))
        return _sweep_reconstructed(Q, p["theta"], dt_, p["dx"], F, flux)
    if scheme == "ee2d_kp07_dimsplit":
        gamma = p["gamma"]
        # Predictor passes gamma in the pressure slot: EE2D_KP07_dimsplit.cu:57-58
        F = lambda q: euler_F(q, gamma)
        return _sweep_reconstructed(Q, p["theta"], dt_, p["dx"], F, lambda a, b: euler_hll_flux(a, b, gamma))
    raise ValueError(scheme)


def _y_sweep(scheme, Q, p, dt_):
    """Y sweep = X sweep on the transposed arrays with the momenta swapped
    (SWE2D_KP07_dimsplit.cu:71-110, EE2D_KP07_dimsplit.cu:73-117)."""
    perm = (0, 2, 1) if len(Q) == 3 else (0, 2, 1, 3)
    Qt = [Q[k].T for k in perm]
    py = dict(p)
    py["dx"] = p["dy"]
    if scheme == "ee2d_kp07_dimsplit":
        gamma = p["gamma"]
        F = lambda q: euler_F(q, gamma)
        res = _sweep_reconstructed(Qt, p["theta"], dt_, p["dy"], F, lambda a, b: euler_cu_flux(a, b, gamma))
    else:
        res = _x_sweep(scheme, Qt, py, dt_)
    out = [None] * len(Q)
    for j, k in enumerate(perm):
        out[k] = np.ascontiguousarray(res[j].T)
    return out


SCHEME_NVARS = {"kp07_dimsplit": 3, "hll2": 3, "waf": 3, "ee2d_kp07_dimsplit": 4}


def substep(scheme, U0, U1, params, dt_, step, bc):
    """One dimensionally split substep from buffers U0 into buffers U1.

    ``dt_`` is the substep length (the wrappers pass ``dt*0.5``,
    KP07_dimsplit.py:108-109).  step 0 sweeps X then Y, otherwise Y then X
    (SWE2D_KP07_dimsplit.cu:161-194).  Only the interior of ``U1`` is
    written (common.h:359-393).  Returns the minimum per-cell CFL dt of the
    result (SWECommon.h:481-532 / EulerCommon.h:26-77).
    """
    ny, nx = U0[0].shape[0] - 4, U0[0].shape[1] - 4
    d = U0[0].dtype.type
    dt_ = d(dt_)
    signs = EULER_SIGNS if scheme == "ee2d_kp07_dimsplit" else SWE_SIGNS
    Q = fill_ghosts(U0, signs, bc, nx, ny)
    if step == 0:
        Q = _x_sweep(scheme, Q, params, dt_)
        Q = _y_sweep(scheme, Q, params, dt_)
    else:
        Q = _y_sweep(scheme, Q, params, dt_)
        Q = _x_sweep(scheme, Q, params, dt_)
    I = (slice(2, ny + 2), slice(2, nx + 2))
    if scheme == "ee2d_kp07_dimsplit":
        g = params["g"]
        if g > d(0.0):  # EE2D_KP07_dimsplit.cu:197-205, 225-233
            rho_v = Q[2][I].copy()
            Q[2][I] = Q[2][I] - g * Q[0][I] * dt_
            Q[3][I] = Q[3][I] - g * rho_v * dt_
    for k in range(len(Q)):
        U1[k][I] = Q[k][I]
    if scheme == "ee2d_kp07_dimsplit":
        c = euler_cfl(Q[0][I], Q[1][I], Q[2][I], params["dx"], params["dy"], params["gamma"])
    else:
        c = swe_cfl(Q[0][I], Q[1][I], Q[2][I], params["dx"], params["dy"], params["g"])
    return d(np.min(c))


def halo_exchange_periodic(U, nx, ny, wrap_x, wrap_y):
    """One-rank halo exchange with itself, MPISimulator.py:247-309, 371-480:
    first north/south (rows [ny, ny+2) -> [0, 2) and [2, 4) -> [ny+2, ny+4),
    interior columns only), then east/west over all ny+4 rows
    (columns [nx, nx+2) -> [0, 2) and [2, 4) -> [nx+2, nx+4))."""
    for a in U:
        if wrap_y:
            south_src = a[ny:ny + 2, 2:nx + 2].copy()
            north_src = a[2:4, 2:nx + 2].copy()
            a[0:2, 2:nx + 2] = south_src
            a[ny + 2:ny + 4, 2:nx + 2] = north_src
        if wrap_x:
            west_src = a[:, nx:nx + 2].copy()
            east_src = a[:, 2:4].copy()
            a[:, nx + 2:nx + 4] = east_src
            a[:, 0:2] = west_src


def halo_copy_periodic_x(U, nx):
    """East/west leg only (the case of a single row-strip rank with periodic x)."""
    ny = U[0].shape[0] - 4
    halo_exchange_periodic(U, nx, ny, True, False)


def pad(field, dtype):
    """Zero-initialised padded buffer with the interior uploaded
    (CudaArray2D.__init__, Common.py:502-531).  Accepts (ny, nx) or padded input."""
    f = np.asarray(field)
    if f.ndim != 2:
        raise ValueError("expected a 2-D field")
    ny, nx = f.shape
    out = np.zeros((ny + 4, nx + 4), dtype=dtype)
    out[2:ny + 2, 2:nx + 2] = f.astype(dtype)
    return out


class OracleSimulator:
    """Stateful restatement of one scheme wrapper plus BaseSimulator.

    GPUSimulators/Simulator.py:176-228 (simulate/step), KP07_dimsplit.py:102-137
    (initial CFL fill, substep dt*0.5, swap, computeDt = 0.5*min), and the same
    in HLL2.py, WAF.py, EE2D_KP07_dimsplit.py (with the buffer swap after every
    substep, the MPISimulator semantics).
    """

    def __init__(self, scheme, fields, nx, ny, dx, dy, g, gamma=None, theta=None,
                 cfl_scale=0.9, bc=None, dtype=np.float32, halo=None,
                 waf_static_dt=False):
        self.scheme = scheme
        self.dtype = np.dtype(dtype)
        d = self.dtype.type
        self.nx, self.ny = int(nx), int(ny)
        if theta is None:
            theta = 1.8 if scheme == "hll2" else 1.3
        self.params = {"dx": d(dx), "dy": d(dy), "g": d(g), "theta": d(theta),
                       "gamma": d(gamma if gamma is not None else 0.0)}
        self.cfl_scale = cfl_scale
        self.bc = pack_bc(REFLECTIVE, REFLECTIVE, REFLECTIVE, REFLECTIVE) if bc is None else int(bc)
        # (wrap_x, wrap_y): one-rank halo exchange after every substep, the
        # MPISimulator path with the periodic edges of `bc` already set to Dirichlet
        self.halo = halo
        self.waf_static_dt = waf_static_dt
        fields = [np.asarray(f) for f in fields]
        self.u0 = [pad(f, dtype) if f.shape == (self.ny, self.nx) else np.array(f, dtype=dtype) for f in fields]
        self.u1 = [np.zeros_like(a) for a in self.u0]
        if halo:
            halo_exchange_periodic(self.u0, self.nx, self.ny, *halo)
        # initial fill, KP07_dimsplit.py:103-106 (host NumPy on the caller's arrays)
        I = (slice(2, self.ny + 2), slice(2, self.nx + 2))
        q = [a[I] for a in self.u0]
        if scheme == "ee2d_kp07_dimsplit":
            c0 = euler_cfl(q[0], q[1], q[2], self.params["dx"], self.params["dy"], self.params["gamma"])
        else:
            c0 = swe_cfl(q[0], q[1], q[2], self.params["dx"], self.params["dy"], self.params["g"])
        self.cfl = d(np.min(c0))
        self.t = 0.0
        self.nt = 0
        self.dt = None

    def substep(self, dt, step_number):
        d = self.dtype.type
        c = substep(self.scheme, self.u0, self.u1, self.params, d(dt) * d(0.5), step_number, self.bc)
        if not (self.scheme == "waf" and self.waf_static_dt):
            self.cfl = c
        self.u0, self.u1 = self.u1, self.u0
        if self.halo:
            halo_exchange_periodic(self.u0, self.nx, self.ny, *self.halo)

    def step(self, dt):
        d = self.dtype.type
        dt = d(dt)
        for i in range(2):
            self.substep(dt, i)
        self.t += float(dt)
        self.nt += 1

    def compute_dt(self):
        """0.5 * min(cfl) (KP07_dimsplit.py:135-137).  The reference era's NumPy
        promoted a 0-d float32 times a Python float to float64, so the product is
        kept in double here; it is rounded to float32 once, at ``current_dt``."""
        return float(self.cfl) * 0.5

    def simulate(self, t, dt=None):
        d = self.dtype.type
        t_end = self.t + t
        update_dt = dt is None
        if dt is not None:
            self.dt = float(dt)
        while self.t < t_end:
            if update_dt and self.nt % 100 == 0:
                self.dt = self.compute_dt() * self.cfl_scale
            current_dt = d(min(self.dt, t_end - self.t))
            if current_dt <= 0.0:
                break
            self.step(current_dt)

    def download(self):
        I = (slice(2, self.ny + 2), slice(2, self.nx + 2))
        return [a[I].copy() for a in self.u0]


def rel_l1(a, b):
    """Relative L1 difference sum|a-b| / sum|b| (b is the reference)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.sum(np.abs(b))
    num = np.sum(np.abs(a - b))
    if den == 0.0:
        return num
    return num / den
