"""CPU pins of the oracle's Euler fluxes (no GPU).

The consistency check flux(Q, Q) = F(Q) cannot see a flux's wave-speed
estimate or its upwind selection: with equal states every dissipation term
multiplies a zero jump.  These tests pin them independently:

* the sonic threshold: a state moving at 1.02 c (c = sqrt(gamma p / rho), the
  sound speed of a perfect gas) is supersonic, so HLL returns F(Q_l) exactly
  and the central-upwind flux returns F(Q_m); at 0.98 c both must differ from
  the upwind flux.  The states are chosen with p / rho far from rho, so a
  sound speed taken from sqrt(gamma rho) (or with a dropped gamma, or from
  p alone) flips at least one of the two verdicts, and so does a swapped
  upwind branch;
* mirror symmetry: reflecting x -> -x swaps the states and flips the normal
  momentum, and the flux must transform as -M F (M = diag(1, -1, 1, 1));
* the Sod shock tube (Toro, "Riemann Solvers and Numerical Methods for Fluid
  Dynamics", 3rd ed., test 1 of section 4.3.3) run through the whole
  EE2D_KP07_dimsplit oracle in x (HLL flux) and in y (central-upwind flux),
  compared with the exact Riemann solution written from Toro chapter 4 (the
  pressure function, its Newton iteration and the sampling of section 4.5).
  The exact solver is first pinned against Toro's published star values
  (Table 4.2: p* = 0.30313, u* = 0.92745, rho*_L = 0.42632, rho*_R = 0.26557).

Covers oracle/fv_oracle.py: euler_hll_flux (EulerCommon.h:117-160),
euler_cu_flux (EulerCommon.h:171-187) and the Euler substep.
"""

import numpy as np
import pytest

from oracle import fv_oracle as O

GAMMA = 1.4


def prim_to_cons(rho, u, v, p, gamma=GAMMA):
    rho, u, v, p = (np.asarray(x, dtype=np.float64) for x in (rho, u, v, p))
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return (rho, rho * u, rho * v, E)


def physical_flux(rho, u, v, p, gamma=GAMMA):
    """x flux of the Euler equations from primitive variables."""
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return (rho * u, rho * u * u + p, rho * u * v, u * (E + p))


def sound_speed(rho, p, gamma=GAMMA):
    return np.sqrt(gamma * p / rho)


# (rho, p) pairs with p / rho far from rho in both directions
SONIC_STATES = [(0.5, 2.0), (2.0, 0.5), (0.2, 1.0), (4.0, 0.3)]


def _flux(name):
    return O.euler_hll_flux if name == "hll" else O.euler_cu_flux


@pytest.mark.parametrize("rho,p", SONIC_STATES)
@pytest.mark.parametrize("name", ["hll", "cu"])
def test_supersonic_right_returns_upwind_flux(name, rho, p):
    """Both states at >= 1.02 c to the right: the flux is the left state's F."""
    c = sound_speed(rho, p)
    u_l = 1.02 * c
    # the right state expands away (faster, same rho and p): no shock, so the
    # HLL star-depth factor q_l stays 1 and S_l = u_l - c_l > 0
    u_r = 1.10 * c
    Ql = prim_to_cons([rho], [u_l], [0.3], [p])
    Qr = prim_to_cons([rho], [u_r], [-0.2], [p])
    out = _flux(name)(Ql, Qr, GAMMA)
    F = physical_flux(rho, u_l, 0.3, p)
    for k in range(4):
        assert np.allclose(out[k], F[k], rtol=1e-13, atol=0.0), (name, k, out[k], F[k])


@pytest.mark.parametrize("rho,p", SONIC_STATES)
@pytest.mark.parametrize("name", ["hll", "cu"])
def test_supersonic_left_returns_downwind_state_flux(name, rho, p):
    """Both states at >= 1.02 c to the left: the flux is the right state's F."""
    c = sound_speed(rho, p)
    Ql = prim_to_cons([rho], [-1.10 * c], [0.1], [p])
    Qr = prim_to_cons([rho], [-1.02 * c], [0.4], [p])
    out = _flux(name)(Ql, Qr, GAMMA)
    F = physical_flux(rho, -1.02 * c, 0.4, p)
    for k in range(4):
        assert np.allclose(out[k], F[k], rtol=1e-13, atol=0.0), (name, k, out[k], F[k])


@pytest.mark.parametrize("rho,p", SONIC_STATES)
@pytest.mark.parametrize("name", ["hll", "cu"])
def test_subsonic_is_not_upwind(name, rho, p):
    """At 0.98 c the left-going wave of the left state moves left, so the flux
    is no longer F(Q_l): it carries the dissipation of the jump."""
    c = sound_speed(rho, p)
    u_l = 0.98 * c
    Ql = prim_to_cons([rho], [u_l], [0.3], [p])
    Qr = prim_to_cons([rho], [1.10 * c], [-0.2], [p])
    out = _flux(name)(Ql, Qr, GAMMA)
    F = physical_flux(rho, u_l, 0.3, p)
    diff = max(abs(float(out[k][0] - F[k])) / (abs(F[k]) + 1e-30) for k in range(4))
    assert diff > 1e-4, (name, diff)


def test_hll_star_flux_closed_form_weights():
    """Subsonic on both sides, equal rho and p, opposite velocities: the HLL
    flux is the Toro (10.21) average with the two-rarefaction speeds, and its
    mass flux vanishes by symmetry (u_l = -u_r)."""
    rho, p = 1.0, 1.0
    c = sound_speed(rho, p)
    u = 0.3 * c
    Ql = prim_to_cons([rho], [-u], [0.0], [p])
    Qr = prim_to_cons([rho], [u], [0.0], [p])
    out = O.euler_hll_flux(Ql, Qr, GAMMA)
    assert abs(out[0][0]) < 1e-14
    # expansion (h_dag < rho), so q = 1: S_l = -u - c, S_r = u + c
    S_l, S_r = -u - c, u + c
    Fl = physical_flux(rho, -u, 0.0, p)
    Fr = physical_flux(rho, u, 0.0, p)
    for k in range(4):
        want = (S_r * Fl[k] - S_l * Fr[k] + S_r * S_l * (Qr[k][0] - Ql[k][0])) / (S_r - S_l)
        assert np.isclose(out[k][0], want, rtol=1e-13, atol=1e-15)


def _random_states(seed, n=400):
    rng = np.random.default_rng(seed)
    rho = 0.2 + 2.0 * rng.random(n)
    p = 0.1 + 2.0 * rng.random(n)
    c = sound_speed(rho, p)
    # velocities from -2c to 2c: every branch of both fluxes is taken
    u = (4.0 * rng.random(n) - 2.0) * c
    v = rng.standard_normal(n)
    return prim_to_cons(rho, u, v, p)


@pytest.mark.parametrize("name", ["hll", "cu"])
def test_mirror_symmetry(name):
    """flux(M Q_r, M Q_l) = -M flux(Q_l, Q_r) with M = diag(1, -1, 1, 1)."""
    Ql = _random_states(11)
    Qr = _random_states(12)

    def mirror(Q):
        return (Q[0], -Q[1], Q[2], Q[3])

    F = _flux(name)(Ql, Qr, GAMMA)
    G = _flux(name)(mirror(Qr), mirror(Ql), GAMMA)
    sign = (-1.0, 1.0, -1.0, -1.0)
    for k in range(4):
        assert np.allclose(G[k], sign[k] * F[k], rtol=1e-12, atol=1e-12), k


def test_every_hll_branch_taken_by_mirror_states():
    """The random states of the mirror test reach all three HLL branches."""
    Ql = _random_states(11)
    Qr = _random_states(12)
    out = O.euler_hll_flux(Ql, Qr, GAMMA)
    Fl = O.euler_F(Ql, O.euler_pressure(Ql, GAMMA))
    Fr = O.euler_F(Qr, O.euler_pressure(Qr, GAMMA))
    left = np.all([out[k] == Fl[k] for k in range(4)], axis=0)
    right = np.all([out[k] == Fr[k] for k in range(4)], axis=0)
    assert left.sum() > 10 and right.sum() > 10 and (~left & ~right).sum() > 10


# --------------------------------------------------------------------------
# exact Riemann solver (Toro chapter 4)
# --------------------------------------------------------------------------

def _f_k(p, rho_k, p_k, c_k, gamma):
    """Toro (4.6)-(4.7): the pressure function of one side and its derivative."""
    if p > p_k:  # shock
        A = 2.0 / ((gamma + 1.0) * rho_k)
        B = (gamma - 1.0) / (gamma + 1.0) * p_k
        s = np.sqrt(A / (p + B))
        return (p - p_k) * s, s * (1.0 - 0.5 * (p - p_k) / (B + p))
    # rarefaction
    e = (gamma - 1.0) / (2.0 * gamma)
    f = 2.0 * c_k / (gamma - 1.0) * ((p / p_k) ** e - 1.0)
    df = 1.0 / (rho_k * c_k) * (p / p_k) ** (-(gamma + 1.0) / (2.0 * gamma))
    return f, df


def exact_star(WL, WR, gamma=GAMMA):
    """p*, u* of the Riemann problem (Toro section 4.3.1, Newton from the
    arithmetic mean pressure)."""
    rl, ul, pl = WL
    rr, ur, pr = WR
    cl, cr = sound_speed(rl, pl, gamma), sound_speed(rr, pr, gamma)
    p = 0.5 * (pl + pr)
    for _ in range(100):
        fl, dfl = _f_k(p, rl, pl, cl, gamma)
        fr, dfr = _f_k(p, rr, pr, cr, gamma)
        p_new = p - (fl + fr + ur - ul) / (dfl + dfr)
        p_new = max(p_new, 1e-12)
        if abs(p_new - p) < 1e-14 * 0.5 * (p_new + p):
            p = p_new
            break
        p = p_new
    fl, _ = _f_k(p, rl, pl, cl, gamma)
    fr, _ = _f_k(p, rr, pr, cr, gamma)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def exact_sample(WL, WR, s, gamma=GAMMA):
    """(rho, u, p) at x/t = s for every entry of s (Toro section 4.5)."""
    rl, ul, pl = WL
    rr, ur, pr = WR
    cl, cr = sound_speed(rl, pl, gamma), sound_speed(rr, pr, gamma)
    ps, us = exact_star(WL, WR, gamma)
    g1 = (gamma - 1.0) / (gamma + 1.0)
    out = np.zeros((3, len(s)))
    for i, x in enumerate(s):
        if x <= us:  # left of the contact
            if ps > pl:  # left shock
                rs = rl * (ps / pl + g1) / (g1 * ps / pl + 1.0)
                sl = ul - cl * np.sqrt((gamma + 1.0) / (2.0 * gamma) * ps / pl + (gamma - 1.0) / (2.0 * gamma))
                out[:, i] = (rl, ul, pl) if x <= sl else (rs, us, ps)
            else:  # left rarefaction
                rs = rl * (ps / pl) ** (1.0 / gamma)
                cs = cl * (ps / pl) ** ((gamma - 1.0) / (2.0 * gamma))
                if x <= ul - cl:
                    out[:, i] = (rl, ul, pl)
                elif x >= us - cs:
                    out[:, i] = (rs, us, ps)
                else:  # inside the fan
                    c = 2.0 / (gamma + 1.0) * (cl + 0.5 * (gamma - 1.0) * (ul - x))
                    u = 2.0 / (gamma + 1.0) * (cl + 0.5 * (gamma - 1.0) * ul + x)
                    r = rl * (c / cl) ** (2.0 / (gamma - 1.0))
                    out[:, i] = (r, u, pl * (c / cl) ** (2.0 * gamma / (gamma - 1.0)))
        else:  # right of the contact
            if ps > pr:  # right shock
                rs = rr * (ps / pr + g1) / (g1 * ps / pr + 1.0)
                sr = ur + cr * np.sqrt((gamma + 1.0) / (2.0 * gamma) * ps / pr + (gamma - 1.0) / (2.0 * gamma))
                out[:, i] = (rr, ur, pr) if x >= sr else (rs, us, ps)
            else:  # right rarefaction
                rs = rr * (ps / pr) ** (1.0 / gamma)
                cs = cr * (ps / pr) ** ((gamma - 1.0) / (2.0 * gamma))
                if x >= ur + cr:
                    out[:, i] = (rr, ur, pr)
                elif x <= us + cs:
                    out[:, i] = (rs, us, ps)
                else:
                    c = 2.0 / (gamma + 1.0) * (cr - 0.5 * (gamma - 1.0) * (ur - x))
                    u = 2.0 / (gamma + 1.0) * (-cr + 0.5 * (gamma - 1.0) * ur + x)
                    r = rr * (c / cr) ** (2.0 / (gamma - 1.0))
                    out[:, i] = (r, u, pr * (c / cr) ** (2.0 * gamma / (gamma - 1.0)))
    return out


SOD_L = (1.0, 0.0, 1.0)
SOD_R = (0.125, 0.0, 0.1)


def test_exact_solver_against_toro_table():
    """Toro Table 4.2, test 1 (Sod): p* = 0.30313, u* = 0.92745; the star
    densities of Table 4.3: 0.42632 (left, behind the rarefaction) and 0.26557
    (right, behind the shock)."""
    ps, us = exact_star(SOD_L, SOD_R)
    assert abs(ps - 0.30313) < 5e-6
    assert abs(us - 0.92745) < 5e-6
    w = exact_sample(SOD_L, SOD_R, np.array([us - 1e-6, us + 1e-6]))
    assert abs(w[0, 0] - 0.42632) < 5e-6
    assert abs(w[0, 1] - 0.26557) < 5e-6
    # the far states are untouched
    w = exact_sample(SOD_L, SOD_R, np.array([-5.0, 5.0]))
    assert np.allclose(w[:, 0], SOD_L) and np.allclose(w[:, 1], SOD_R)


def _sod_run(n, axis, t_end=0.2, width=4):
    """The Sod tube on [0, 1] with n cells along `axis` ('x' or 'y') and
    `width` uniform cells across, reflective walls (the waves stay inside by
    t = 0.2), no gravity, float64 oracle, fixed dt = 0.3 dx (the largest wave
    speed is below 1.8, so the Courant number stays below 0.55 per step and
    0.27 per substep)."""
    dx = 1.0 / n
    xc = (np.arange(n) + 0.5) * dx
    left = xc < 0.5
    rho1 = np.where(left, SOD_L[0], SOD_R[0])
    p1 = np.where(left, SOD_L[2], SOD_R[2])
    z = np.zeros(n)
    Q = prim_to_cons(rho1, z, z, p1)
    if axis == "x":
        fields = [np.tile(q, (width, 1)) for q in Q]
        nx, ny = n, width
    else:  # along y: rho_v carries the normal momentum
        fields = [np.tile(q[:, None], (1, width)) for q in Q]
        nx, ny = width, n
    sim = O.OracleSimulator("ee2d_kp07_dimsplit", fields, nx, ny, dx, dx, 0.0, gamma=GAMMA, dtype=np.float64)
    dt = 0.3 * dx
    steps = int(round(t_end / dt))
    for _ in range(steps):
        sim.step(dt)
    rho, ru, rv, E = sim.download()
    if axis == "x":
        rho, mom, tang, E = rho[width // 2], ru[width // 2], rv[width // 2], E[width // 2]
    else:
        rho, mom, tang, E = rho[:, width // 2], rv[:, width // 2], ru[:, width // 2], E[:, width // 2]
    u = mom / rho
    p = (GAMMA - 1.0) * (E - 0.5 * mom * u)
    ex = exact_sample(SOD_L, SOD_R, (xc - 0.5) / (steps * dt))
    return np.array([rho, u, p]), ex, tang, (rho, mom, E, steps * dt)


@pytest.mark.parametrize("axis", ["x", "y"])
def test_sod_against_exact_riemann_solution(axis):
    """x runs the HLL flux, y the central-upwind flux.  The L1 error of rho,
    u and p against the exact solution is below a stated bound at 400 cells
    and falls by at least a third with every doubling from 100 to 400 cells;
    the tangential momentum stays exactly zero."""
    errs = []
    for n in (100, 200, 400):
        num, ex, tang, _ = _sod_run(n, axis)
        assert np.all(np.isfinite(num))
        assert np.all(tang == 0.0)
        assert np.all(num[0] > 0) and np.all(num[2] > 0)
        errs.append(np.mean(np.abs(num - ex), axis=1))  # L1 over [0, 1] (dx = 1/n)
    errs = np.array(errs)
    # stated bounds at 400 cells (rho, u, p); measured 0.00151, 0.00273,
    # 0.00093 in x and 0.00152, 0.00271, 0.00094 in y
    assert np.all(errs[2] < np.array([0.002, 0.0035, 0.0013])), errs
    # convergence under refinement (measured ratios 0.43-0.58 per doubling)
    assert np.all(errs[1] < 0.65 * errs[0]) and np.all(errs[2] < 0.65 * errs[1]), errs


def test_sod_conserves_mass_momentum_energy_exactly_enough():
    """Reflective walls the waves never reach: the totals of rho and E do not
    change and the momentum equals the wall-pressure impulse, zero here."""
    _, _, _, (rho, mom, E, t) = _sod_run(200, "x")
    dx = 1.0 / 200
    mass0 = 0.5 * SOD_L[0] + 0.5 * SOD_R[0]
    energy0 = 0.5 * SOD_L[2] / (GAMMA - 1.0) + 0.5 * SOD_R[2] / (GAMMA - 1.0)
    assert abs(rho.sum() * dx - mass0) < 1e-12
    assert abs(E.sum() * dx - energy0) < 1e-12
    # the pressure at both walls is unchanged (1.0 and 0.1), so the total
    # momentum grows by (p_left - p_right) t
    assert abs(mom.sum() * dx - (SOD_L[2] - SOD_R[2]) * t) < 1e-10
