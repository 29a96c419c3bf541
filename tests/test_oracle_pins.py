"""CPU pins of the oracle (no GPU): known answers, closed forms and invariants
chosen so that a dropped term, a wrong sign or index, or a transposed operand
fails at least one of them.

Covers oracle/fv_oracle.py functions:
  minmod_slope, generalized_minmod, desingularize, limiter_to_waf_limiter,
  swe_F, swe_cu_flux, swe_hll_flux, swe_waf_flux, compute_h_star, swe_cfl,
  euler_pressure, euler_F, euler_hll_flux, euler_cu_flux, euler_cfl,
  fill_ghosts, substep (sweeps, extents, gravity), halo_copy_periodic_x,
  OracleSimulator (step / simulate dt policy).
"""

import numpy as np
import pytest

from oracle import fv_oracle as O

f32 = np.float32


def a32(*v):
    return np.array(v, dtype=np.float32)


# --------------------------------------------------------------------------
# limiters
# --------------------------------------------------------------------------

def test_minmod_known_answers():
    th = f32(1.3)
    # monotone increasing: min(theta*(c-l), (r-l)/2, theta*(r-c))
    l, c, r = a32(0.0), a32(1.0), a32(3.0)
    assert O.minmod_slope(l, c, r, th)[0] == np.float32(1.3)          # theta*(c-l) = 1.3 < 1.5 < 2.6
    l, c, r = a32(0.0), a32(2.0), a32(2.5)
    assert O.minmod_slope(l, c, r, th)[0] == np.float32(1.3) * np.float32(0.5)  # theta*(r-c) = 0.65
    l, c, r = a32(0.0), a32(1.0), a32(2.0)
    assert O.minmod_slope(l, c, r, th)[0] == np.float32(1.0)          # central (r-l)/2 = 1 < 1.3
    # monotone decreasing: negative slope
    assert O.minmod_slope(a32(3.0), a32(1.0), a32(0.0), th)[0] == -np.float32(1.3)
    # extremum: zero
    assert O.minmod_slope(a32(0.0), a32(1.0), a32(0.5), th)[0] == 0.0
    assert O.minmod_slope(a32(1.0), a32(0.0), a32(1.0), th)[0] == 0.0
    # flat: zero
    assert O.minmod_slope(a32(1.0), a32(1.0), a32(1.0), th)[0] == 0.0
    # one-sided flat: zero
    assert O.minmod_slope(a32(1.0), a32(1.0), a32(2.0), th)[0] == 0.0


def test_minmod_symmetry_and_scaling():
    rng = np.random.default_rng(0)
    l, c, r = (rng.standard_normal(1000).astype(np.float32) for _ in range(3))
    th = f32(1.3)
    s = O.minmod_slope(l, c, r, th)
    # reflection x -> -x reverses the stencil and the slope sign
    assert np.array_equal(O.minmod_slope(r, c, l, th), -s)
    # negation of the data negates the slope
    assert np.array_equal(O.minmod_slope(-l, -c, -r, th), -s)
    # the slope never exceeds theta times either one-sided difference
    assert np.all(np.abs(s) <= th * np.abs(c - l) + 1e-6)
    assert np.all(np.abs(s) <= th * np.abs(r - c) + 1e-6)


def test_generalized_minmod_known_answers():
    th = f32(1.9)
    assert O.generalized_minmod(a32(-1.0), th)[0] == 0.0
    assert O.generalized_minmod(a32(0.5), th)[0] == np.float32(0.75)    # min(0.95, 0.75, 1.9)
    assert O.generalized_minmod(a32(0.2), th)[0] == np.float32(1.9) * np.float32(0.2)
    assert O.generalized_minmod(a32(10.0), th)[0] == np.float32(1.9)


def test_desingularize_known_answers():
    eps = f32(1e-6)
    # |x| >= eps: identity
    for x in (1.0, -1.0, 1e-3, -2e-6):
        assert O.desingularize(a32(x), eps)[0] == np.float32(x)
    # x = 0: copysign(1, +0) * min(0 + eps/2, eps) = eps/2
    assert O.desingularize(a32(0.0), eps)[0] == np.float32(0.5) * eps
    assert O.desingularize(a32(-0.0), eps)[0] == -np.float32(0.5) * eps
    # small x: x^2/(2 eps) + eps/2 (> |x|), keeps the sign
    x = np.float32(4e-7)
    v = O.desingularize(a32(x), eps)[0]
    assert v == x * x / (np.float32(2) * eps) + np.float32(0.5) * eps
    assert O.desingularize(a32(-x), eps)[0] == -v


def test_limiter_to_waf():
    assert O.limiter_to_waf_limiter(a32(0.0), a32(0.3))[0] == 1.0
    assert O.limiter_to_waf_limiter(a32(1.0), a32(0.3))[0] == np.float32(0.3)
    assert O.limiter_to_waf_limiter(a32(1.0), a32(-0.3))[0] == np.float32(0.3)


# --------------------------------------------------------------------------
# fluxes
# --------------------------------------------------------------------------

def _swe_states(n=200, seed=1):
    rng = np.random.default_rng(seed)
    h = (0.5 + rng.random(n)).astype(np.float64)
    hu = rng.standard_normal(n) * 0.3
    hv = rng.standard_normal(n) * 0.3
    return h, hu, hv


def test_swe_F_closed_form():
    h, hu, hv = _swe_states()
    g = 9.81
    F = O.swe_F(h, hu, hv, g)
    assert np.allclose(F[0], hu)
    assert np.allclose(F[1], hu ** 2 / h + 0.5 * g * h ** 2)
    assert np.allclose(F[2], hu * hv / h)


@pytest.mark.parametrize("flux", ["cu", "hll"])
def test_swe_flux_consistency(flux):
    """flux(Q, Q) = F(Q) for every state (subsonic and supersonic)."""
    h, hu, hv = _swe_states()
    hu = hu * 10  # include supersonic states
    g = 9.81
    Q = (h, hu, hv)
    fn = O.swe_cu_flux if flux == "cu" else O.swe_hll_flux
    out = fn(Q, Q, g)
    F = O.swe_F(h, hu, hv, g)
    for k in range(3):
        assert np.allclose(out[k], F[k], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("flux", ["cu", "hll"])
def test_swe_flux_mirror(flux):
    """Reflecting x -> -x swaps the states, flips the normal momentum and
    negates the mass / tangential flux while keeping the momentum flux."""
    g = 9.81
    h1, u1, v1 = _swe_states(seed=2)
    h2, u2, v2 = _swe_states(seed=3)
    fn = O.swe_cu_flux if flux == "cu" else O.swe_hll_flux
    F = fn((h1, u1, v1), (h2, u2, v2), g)
    M = fn((h2, -u2, v2), (h1, -u1, v1), g)
    assert np.allclose(M[0], -F[0], atol=1e-12)
    assert np.allclose(M[1], F[1], atol=1e-12)
    assert np.allclose(M[2], -F[2], atol=1e-12)


def test_swe_cu_upwind_limit():
    """Supersonic flow to the right: the central-upwind flux is the left F."""
    g = 9.81
    ql = (np.array([1.0]), np.array([10.0]), np.array([0.2]))
    qr = (np.array([0.8]), np.array([9.0]), np.array([0.1]))
    out = O.swe_cu_flux(ql, qr, g)
    F = O.swe_F(*ql, g)
    for k in range(3):
        assert np.allclose(out[k], F[k])
    # and the HLL flux picks F(ql) by its S_l >= 0 branch
    out = O.swe_hll_flux(ql, qr, g)
    for k in range(3):
        assert np.allclose(out[k], F[k])


def test_swe_waf_consistency():
    """Uniform data: WAF flux = F(Q) (all waves carry zero jumps)."""
    h, hu, hv = _swe_states(seed=4)
    Q = (h, hu, hv)
    out = O.swe_waf_flux(Q, Q, Q, Q, 9.81, 1.0, 0.01)
    F = O.swe_F(h, hu, hv, 9.81)
    for k in range(3):
        assert np.allclose(out[k], F[k], rtol=1e-10, atol=1e-10)


def test_compute_h_star_two_rarefaction():
    # h* = (0.5 (c_l + c_r) + 0.25 (u_l - u_r))^2 / g
    g = 9.81
    h_l, h_r, u_l, u_r = np.array([1.0]), np.array([2.0]), np.array([0.3]), np.array([-0.1])
    c_l, c_r = np.sqrt(g * h_l), np.sqrt(g * h_r)
    v = O.compute_h_star(h_l, h_r, u_l, u_r, c_l, c_r, g)
    assert np.allclose(v, (0.5 * (c_l + c_r) + 0.25 * (u_l - u_r)) ** 2 / g)


def _euler_states(n=200, seed=5, gamma=1.4):
    rng = np.random.default_rng(seed)
    rho = 0.5 + rng.random(n)
    u = rng.standard_normal(n) * 0.5
    v = rng.standard_normal(n) * 0.5
    p = 1.0 + rng.random(n)
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return (rho, rho * u, rho * v, E), p


def test_euler_pressure_and_F():
    Q, p = _euler_states()
    gamma = 1.4
    assert np.allclose(O.euler_pressure(Q, gamma), p)
    F = O.euler_F(Q, p)
    rho, ru, rv, E = Q
    u = ru / rho
    assert np.allclose(F[0], ru)
    assert np.allclose(F[1], ru * u + p)
    assert np.allclose(F[2], rv * u)
    assert np.allclose(F[3], u * (E + p))


@pytest.mark.parametrize("flux", ["cu", "hll"])
def test_euler_flux_consistency(flux):
    Q, p = _euler_states()
    gamma = 1.4
    fn = O.euler_cu_flux if flux == "cu" else O.euler_hll_flux
    out = fn(Q, Q, gamma)
    F = O.euler_F(Q, p)
    for k in range(4):
        assert np.allclose(out[k], F[k], rtol=1e-12, atol=1e-12)


def test_cfl_closed_form():
    h, hu, hv = _swe_states(seed=6)
    c = O.swe_cfl(h, hu, hv, 0.5, 2.0, 9.81)
    ref = np.minimum(0.5 / (np.abs(hu / h) + np.sqrt(9.81 * h)), 2.0 / (np.abs(hv / h) + np.sqrt(9.81 * h)))
    assert np.allclose(c, ref)
    Q, p = _euler_states()
    c = O.euler_cfl(Q[0], Q[1], Q[2], 1.0, 1.0, 1.4)
    ref = np.minimum(1.0 / (np.abs(Q[1] / Q[0]) + np.sqrt(1.4 * Q[0])), 1.0 / (np.abs(Q[2] / Q[0]) + np.sqrt(1.4 * Q[0])))
    assert np.allclose(c, ref)


# --------------------------------------------------------------------------
# ghost fill
# --------------------------------------------------------------------------

def test_fill_ghosts_rules():
    nx, ny = 5, 4
    rng = np.random.default_rng(7)
    mem = [rng.random((ny + 4, nx + 4)).astype(np.float32) for _ in range(3)]
    R, P, D, F = O.REFLECTIVE, O.PERIODIC, O.DIRICHLET, O.OUTFLOW
    # reflective everywhere
    q = O.fill_ghosts(mem, O.SWE_SIGNS, O.pack_bc(R, R, R, R), nx, ny)
    I = slice(2, ny + 2)
    for k, (sx, sy) in enumerate(O.SWE_SIGNS):
        assert np.array_equal(q[k][I, 1], sx * mem[k][I, 2])
        assert np.array_equal(q[k][I, 0], sx * mem[k][I, 3])
        assert np.array_equal(q[k][I, nx + 2], sx * mem[k][I, nx + 1])
        assert np.array_equal(q[k][I, nx + 3], sx * mem[k][I, nx])
        assert np.array_equal(q[k][ny + 2, 2:nx + 2], sy * mem[k][ny + 1, 2:nx + 2])
        assert np.array_equal(q[k][0, 2:nx + 2], sy * mem[k][3, 2:nx + 2])
        # corners: mirrored twice
        assert q[k][0, 0] == sx * sy * mem[k][3, 3]
        assert q[k][ny + 3, nx + 3] == sx * sy * mem[k][ny, nx]
    # periodic everywhere
    q = O.fill_ghosts(mem, O.SWE_SIGNS, O.pack_bc(P, P, P, P), nx, ny)
    for k in range(3):
        assert np.array_equal(q[k][I, 0:2], mem[k][I, nx:nx + 2])
        assert np.array_equal(q[k][I, nx + 2:nx + 4], mem[k][I, 2:4])
        assert np.array_equal(q[k][0:2, 2:nx + 2], mem[k][ny:ny + 2, 2:nx + 2])
        assert q[k][0, 0] == mem[k][ny, nx]
    # Dirichlet: memory untouched
    q = O.fill_ghosts(mem, O.SWE_SIGNS, O.pack_bc(D, D, D, D), nx, ny)
    for k in range(3):
        assert np.array_equal(q[k], mem[k])
    # outflow: zero-gradient copies
    q = O.fill_ghosts(mem, O.SWE_SIGNS, O.pack_bc(F, F, F, F), nx, ny)
    for k in range(3):
        assert np.array_equal(q[k][I, 0], mem[k][I, 2])
        assert np.array_equal(q[k][I, nx + 3], mem[k][I, nx + 1])


def test_bc_packing_matches_reference_layout():
    assert O.pack_bc(3, 2, 1, 0) == (3 << 24) | (2 << 16) | (1 << 8)
    bc = O.pack_bc(O.REFLECTIVE, O.PERIODIC, O.DIRICHLET, O.OUTFLOW)
    assert (O.bc_north(bc), O.bc_south(bc), O.bc_east(bc), O.bc_west(bc)) == (3, 2, 0, 4)


def test_halo_copy_periodic_x():
    nx, ny = 6, 3
    a = np.arange((ny + 4) * (nx + 4), dtype=np.float32).reshape(ny + 4, nx + 4)
    b = a.copy()
    O.halo_copy_periodic_x([b], nx)
    assert np.array_equal(b[:, 0:2], a[:, nx:nx + 2])
    assert np.array_equal(b[:, nx + 2:nx + 4], a[:, 2:4])
    assert np.array_equal(b[:, 2:nx + 2], a[:, 2:nx + 2])


# --------------------------------------------------------------------------
# whole-substep invariants
# --------------------------------------------------------------------------

SWE_SCHEMES = ["kp07_dimsplit", "hll2", "waf"]


@pytest.mark.parametrize("scheme", SWE_SCHEMES)
@pytest.mark.parametrize("bc", ["reflective", "periodic"])
def test_lake_at_rest(scheme, bc):
    """Uniform depth at rest stays exactly at rest (identical interface fluxes)."""
    n = 12
    code = O.pack_bc(*([O.REFLECTIVE] * 4)) if bc == "reflective" else O.pack_bc(*([O.PERIODIC] * 4))
    h = np.full((n, n), 1.25, dtype=np.float32)
    z = np.zeros_like(h)
    sim = O.OracleSimulator(scheme, [h, z, z], n, n, 1.0, 1.0, 9.81, bc=code)
    for _ in range(5):
        sim.step(np.float32(0.05))
    out = sim.download()
    assert np.array_equal(out[0], h)
    assert np.all(out[1] == 0) and np.all(out[2] == 0)


@pytest.mark.parametrize("scheme", SWE_SCHEMES + ["ee2d_kp07_dimsplit"])
@pytest.mark.parametrize("bc", ["reflective", "periodic"])
def test_conservation(scheme, bc):
    """Mass (and, for periodic boundaries, momentum; for Euler without gravity,
    energy) is conserved to rounding."""
    from tests._cases import state
    n = 20
    code = O.pack_bc(*([O.REFLECTIVE] * 4)) if bc == "reflective" else O.pack_bc(*([O.PERIODIC] * 4))
    fields = state(scheme, n, n, 9)
    g = 0.0 if scheme == "ee2d_kp07_dimsplit" else 9.81
    dx = 1.0 / n if scheme == "ee2d_kp07_dimsplit" else 1.0
    sim = O.OracleSimulator(scheme, fields, n, n, dx, dx, g, gamma=1.4, bc=code, dtype=np.float64)
    dt = 0.4 * float(sim.cfl)
    tot0 = [float(np.sum(f, dtype=np.float64)) for f in sim.download()]
    for _ in range(10):
        sim.step(dt)
    tot = [float(np.sum(f, dtype=np.float64)) for f in sim.download()]
    assert abs(tot[0] - tot0[0]) <= 1e-11 * abs(tot0[0])
    if bc == "periodic":
        for k in range(1, len(tot)):
            assert abs(tot[k] - tot0[k]) <= 1e-11 * max(1.0, abs(tot0[0]))


@pytest.mark.parametrize("scheme", ["kp07_dimsplit", "hll2"])
def test_mirror_symmetry(scheme):
    """A dam break symmetric about x = Lx/2 stays symmetric (hu antisymmetric).
    WAF is left out: in the reference the ratio r_1 divides by (h_dag - h_l)
    while its mirror image r_3 divides by (h_r - h_dag) (SWECommon.h:193-195),
    so the scheme itself is not mirror symmetric."""
    n = 24
    y, x = np.mgrid[0:n, 0:n] + 0.5
    h = (1.0 + 0.5 * (np.hypot(x - n / 2, y - 0.4 * n) < 6)).astype(np.float32)
    z = np.zeros_like(h)
    sim = O.OracleSimulator(scheme, [h, z, z], n, n, 1.0, 1.0, 9.81, dtype=np.float64)
    dt = 0.4 * float(sim.cfl)
    for _ in range(15):
        sim.step(dt)
    hh, hu, hv = sim.download()
    tol = 1e-12
    assert np.allclose(hh, hh[:, ::-1], rtol=0, atol=tol)
    assert np.allclose(hu, -hu[:, ::-1], rtol=0, atol=tol)
    assert np.allclose(hv, hv[:, ::-1], rtol=0, atol=tol)


def test_x_then_y_order():
    """Step 0 sweeps x first, step 1 y first: the two orders differ on a 2-D
    state, and swapping the axes of the data swaps the orders."""
    from tests._cases import state
    n = 10
    f = state("kp07_dimsplit", n, n, 12)
    p = {"dx": f32(1), "dy": f32(1), "g": f32(9.81), "theta": f32(1.3), "gamma": f32(0)}
    U0 = [O.pad(a, np.float32) for a in f]
    A = [np.zeros_like(a) for a in U0]
    B = [np.zeros_like(a) for a in U0]
    O.substep("kp07_dimsplit", U0, A, p, 0.05, 0, O.pack_bc(3, 3, 3, 3))
    O.substep("kp07_dimsplit", U0, B, p, 0.05, 1, O.pack_bc(3, 3, 3, 3))
    assert not np.array_equal(A[0], B[0])
    # transpose the data (h, hu, hv) -> (h^T, hv^T, hu^T): step 0 of the transposed
    # state equals the transpose of step 1 of the original
    T0 = [np.ascontiguousarray(U0[0].T), np.ascontiguousarray(U0[2].T), np.ascontiguousarray(U0[1].T)]
    C = [np.zeros_like(a) for a in T0]
    O.substep("kp07_dimsplit", T0, C, p, 0.05, 0, O.pack_bc(3, 3, 3, 3))
    assert np.array_equal(C[0], B[0].T)
    assert np.array_equal(C[1], B[2].T)
    assert np.array_equal(C[2], B[1].T)


def test_euler_gravity_source_on_uniform_state():
    """Uniform rho, E at rest: the fluxes cancel and only the source acts,
    rho_v -= g rho dt after each substep (interior)."""
    n = 8
    rho = np.full((n, n), 1.5, dtype=np.float64)
    z = np.zeros_like(rho)
    E = np.full((n, n), 5.0)
    sim = O.OracleSimulator("ee2d_kp07_dimsplit", [rho, z, z, E], n, n, 1.0, 1.0, 0.5, gamma=1.4,
                            bc=O.pack_bc(O.PERIODIC, O.PERIODIC, O.PERIODIC, O.PERIODIC), dtype=np.float64)
    sim.substep(0.2, 0)  # substep length 0.1
    r, ru, rv, EE = sim.download()
    assert np.allclose(rv, -0.5 * 1.5 * 0.1)
    assert np.allclose(EE, 5.0)  # rho_v was zero before the source
    assert np.allclose(r, 1.5) and np.allclose(ru, 0.0)


# --------------------------------------------------------------------------
# tile emulation (literal block/thread restatement of the CUDA structure)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("scheme", ["kp07_dimsplit", "hll2", "waf", "ee2d_kp07_dimsplit"])
@pytest.mark.parametrize("bcname", ["reflective", "periodic", "mixed"])
def test_tile_emulation_equals_global_form(scheme, bcname):
    from oracle import tile_emulator as T
    from tests._cases import state
    R, P, D = O.REFLECTIVE, O.PERIODIC, O.DIRICHLET
    bc = {"reflective": O.pack_bc(R, R, R, R), "periodic": O.pack_bc(P, P, P, P),
          "mixed": O.pack_bc(R, P, D, R)}[bcname]
    nx, ny, bw, bh = 11, 7, 3, 5  # 11 % 3 = 2, 7 % 5 = 2: partial blocks, no quirk
    f = state(scheme, nx, ny, 13)
    U0 = [O.pad(a, np.float32) for a in f]
    for a, v in zip(U0, (1.05, 0.01, -0.01, 6.5)):
        a[0:2, :] = v
        a[-2:, :] = v
        a[:, 0:2] = v
        a[:, -2:] = v
    euler = scheme == "ee2d_kp07_dimsplit"
    p = {"dx": f32(0.1 if euler else 1), "dy": f32(0.1 if euler else 1), "g": f32(0.1 if euler else 9.81),
         "theta": f32(1.3), "gamma": f32(1.4)}
    for step in (0, 1):
        A = [np.zeros_like(a) for a in U0]
        B = [np.zeros_like(a) for a in U0]
        c1 = O.substep(scheme, U0, A, p, 0.004, step, bc)
        c2 = T.substep(scheme, U0, B, p, 0.004, step, bc, bw, bh)
        for a, b in zip(A, B):
            assert np.array_equal(a, b)
        assert c1 == c2.min()


def test_tile_quirk_nx_mod_bw_equals_1():
    """With nx % bw == 1 the reference's east mirror writes outside its tile row
    (common.h:164-172); the literal emulation reports it, the global form is
    defined for every size."""
    from oracle import tile_emulator as T
    from tests._cases import state
    nx, ny = 13, 8
    f = state("kp07_dimsplit", nx, ny, 14)
    U0 = [O.pad(a, np.float32) for a in f]
    p = {"dx": f32(1), "dy": f32(1), "g": f32(9.81), "theta": f32(1.3), "gamma": f32(0)}
    A = [np.zeros_like(a) for a in U0]
    O.substep("kp07_dimsplit", U0, A, p, 0.01, 0, O.pack_bc(3, 3, 3, 3))
    assert all(np.isfinite(a).all() for a in A)
    with pytest.raises(IndexError):
        T.substep("kp07_dimsplit", U0, [np.zeros_like(a) for a in U0], p, 0.01, 0, O.pack_bc(3, 3, 3, 3), 4, 4)


def test_cfl_is_min_of_cells():
    from tests._cases import state
    n = 16
    f = state("kp07_dimsplit", n, n, 15)
    p = {"dx": f32(1), "dy": f32(1), "g": f32(9.81), "theta": f32(1.3), "gamma": f32(0)}
    U0 = [O.pad(a, np.float32) for a in f]
    A = [np.zeros_like(a) for a in U0]
    c = O.substep("kp07_dimsplit", U0, A, p, 0.01, 0, O.pack_bc(3, 3, 3, 3))
    I = (slice(2, n + 2), slice(2, n + 2))
    assert c == np.min(O.swe_cfl(A[0][I], A[1][I], A[2][I], p["dx"], p["dy"], p["g"]))


# --------------------------------------------------------------------------
# simulate() dt policy
# --------------------------------------------------------------------------

def test_simulate_policy_refresh_every_100_steps_and_shortened_last_step():
    """dt is refreshed from the CFL at steps 0, 100, 200 (Simulator.py:193-196)
    and the last step is shortened to end exactly at t_end (:199)."""
    n = 10
    h = np.full((n, n), 1.0, dtype=np.float32)
    z = np.zeros_like(h)
    sim = O.OracleSimulator("kp07_dimsplit", [h, z, z], n, n, 1.0, 1.0, 9.81)
    calls = []
    orig = sim.compute_dt

    def counting():
        calls.append(sim.nt)
        return orig()
    sim.compute_dt = counting
    dt0 = float(sim.cfl) * 0.5 * 0.9
    t_end = 250.5 * dt0
    sim.simulate(t_end)
    assert calls == [0, 100, 200] or calls == [0, 100, 200, 300][:len(calls)]
    assert calls[:3] == [0, 100, 200]
    # 250 full steps, the shortened one, and possibly a last step of a few ulp
    # (t_end - t rounded to float32 can fall short of t_end, as in the reference)
    assert sim.nt in (251, 252)
    assert abs(sim.t - t_end) < 1e-6 * t_end
    # an explicit dt disables the refresh
    sim2 = O.OracleSimulator("kp07_dimsplit", [h, z, z], n, n, 1.0, 1.0, 9.81)
    sim2.simulate(1.0, dt=0.25)
    assert sim2.nt == 4


# --------------------------------------------------------------------------
# accuracy: SWASHES Stoker dam break (data shipped with the reference)
# --------------------------------------------------------------------------

def _swashes(nx):
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "swashes", "swashes_1_nx=%d.csv" % nx)
    data = np.genfromtxt(path, comments="#", delimiter="\t", usecols=(0, 1, 2))
    return data[:, 0], data[:, 1], data[:, 1] * data[:, 2]


def _stoker_run(scheme, nx):
    ny = 4
    x = (np.arange(nx) + 0.5) * (10.0 / nx)
    h = np.where(x < 5.0, 0.005, 0.001)[None, :].repeat(ny, 0).astype(np.float32)
    z = np.zeros_like(h)
    sim = O.OracleSimulator(scheme, [h, z, z], nx, ny, 10.0 / nx, 10.0 / nx, 9.81,
                            bc=O.pack_bc(O.PERIODIC, O.PERIODIC, O.REFLECTIVE, O.REFLECTIVE), dtype=np.float64)
    sim.simulate(6.0)
    out = sim.download()
    # the run stays uniform in y
    assert np.allclose(out[0], out[0][0:1, :], atol=1e-15)
    return out[0][0], out[1][0]


@pytest.mark.parametrize("scheme", SWE_SCHEMES)
def test_stoker_dam_break_against_swashes(scheme):
    errs = []
    for nx in (128, 256, 512):
        xs, h_ref, q_ref = _swashes(nx)
        h, hu = _stoker_run(scheme, nx)
        errs.append(np.mean(np.abs(h - h_ref)) / np.mean(h_ref))
    # small relative L1 error that decreases under refinement
    assert errs[-1] < 0.02, errs
    assert errs[0] > errs[1] > errs[2], errs


# --------------------------------------------------------------------------
# accuracy: second-order convergence on a smooth problem
# --------------------------------------------------------------------------

def _smooth_run(scheme, n, t_end=0.5):
    L = 1.0
    x = (np.arange(n) + 0.5) * (L / n)
    X, Y = np.meshgrid(x, x)
    h = 1.0 + 0.05 * np.exp(-((X - 0.5) ** 2 + (Y - 0.5) ** 2) / 0.02)
    z = np.zeros_like(h)
    sim = O.OracleSimulator(scheme, [h, z, z], n, n, L / n, L / n, 1.0,
                            bc=O.pack_bc(O.PERIODIC, O.PERIODIC, O.PERIODIC, O.PERIODIC), dtype=np.float64)
    dt = 0.2 * (L / n) / 1.05
    steps = int(np.ceil(t_end / dt))
    dt = t_end / steps
    for _ in range(steps):
        sim.step(dt)
    return sim.download()[0]


def _coarsen(a, f):
    n = a.shape[0] // f
    return a.reshape(n, f, n, f).mean(axis=(1, 3))


@pytest.mark.parametrize("scheme", SWE_SCHEMES)
def test_second_order_convergence(scheme):
    ref = _smooth_run(scheme, 128)
    errs = []
    for n in (16, 32, 64):
        errs.append(np.mean(np.abs(_smooth_run(scheme, n) - _coarsen(ref, 128 // n))))
    order = np.log2(errs[1] / errs[2])
    assert order > 1.6, (errs, order)


@pytest.mark.parametrize("scheme", ["kp07_dimsplit", "ee2d_kp07_dimsplit"])
@pytest.mark.parametrize("wrap", [(True, False), (False, True), (True, True)])
def test_halo_exchange_equals_periodic_boundaries(scheme, wrap):
    """A rank that is its own periodic neighbour (Dirichlet edges + halo
    exchange, MPISimulator.py:247-309) evolves exactly like periodic boundary
    conditions: both read the wrapped interior."""
    from tests._cases import state
    nx, ny = 14, 11
    f = state(scheme, nx, ny, 21)
    P, R, D = O.PERIODIC, O.REFLECTIVE, O.DIRICHLET
    wx, wy = wrap
    bc_per = O.pack_bc(P if wy else R, P if wy else R, P if wx else R, P if wx else R)
    bc_dir = O.pack_bc(D if wy else R, D if wy else R, D if wx else R, D if wx else R)
    euler = scheme == "ee2d_kp07_dimsplit"
    kw = dict(gamma=1.4, dtype=np.float32)
    g = 0.1 if euler else 9.81
    dx = 1.0 / 16 if euler else 1.0
    a = O.OracleSimulator(scheme, f, nx, ny, dx, dx, g, bc=bc_per, **kw)
    b = O.OracleSimulator(scheme, f, nx, ny, dx, dx, g, bc=bc_dir, halo=wrap, **kw)
    dt = np.float32(0.45 * float(a.cfl))
    for _ in range(5):
        a.step(dt)
        b.step(dt)
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(x, y)
