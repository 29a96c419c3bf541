"""Pins of the unsplit / first-order oracle (oracle/fv_oracle_unsplit.py): the
ghost fill against the two-ghost-cell oracle and a scalar restatement, flux
consistency and closed forms, lake at rest, conservation, mirror symmetry,
1-D reductions against independent scalar loops, the SWASHES Stoker fixtures
and the observed order of accuracy."""

import numpy as np
import pytest

from oracle import fv_oracle as O
from oracle import fv_oracle_unsplit as U

SCHEMES = ["lxf", "hll", "force", "kp07"]
R, P, D, F = O.REFLECTIVE, O.PERIODIC, O.DIRICHLET, O.OUTFLOW


def _state(nx, ny, seed, dtype=np.float64):
    rng = np.random.default_rng(seed)
    h = 1.0 + 0.3 * rng.random((ny, nx))
    hu = 0.2 * rng.standard_normal((ny, nx))
    hv = 0.2 * rng.standard_normal((ny, nx))
    return [a.astype(dtype) for a in (h, hu, hv)]


def _run(scheme, fields, bc, steps, dt, dtype=np.float64, nx=None, ny=None, order=2, dx=1.0, dy=1.0, g=9.81):
    ny_, nx_ = fields[0].shape
    sim = U.OracleUnsplitSimulator(scheme, fields, nx_, ny_, dx, dy, g, order=order, bc=bc, dtype=dtype)
    for _ in range(steps):
        sim.step(dtype(dt) if dtype is np.float32 else dt)
    return sim


# --------------------------------------------------------------------------
# ghost fill
# --------------------------------------------------------------------------

@pytest.mark.parametrize("bc", [O.pack_bc(R, R, R, R), O.pack_bc(P, P, P, P), O.pack_bc(R, P, D, F),
                                O.pack_bc(F, F, F, F), O.pack_bc(P, R, F, P)])
def test_ghost_fill_gc2_equals_dimsplit_oracle(bc):
    nx, ny = 9, 7
    rng = np.random.default_rng(1)
    mem = [rng.standard_normal((ny + 4, nx + 4)) for _ in range(3)]
    a = O.fill_ghosts(mem, O.SWE_SIGNS, bc, nx, ny)
    b = U.fill_ghosts_gc(mem, O.SWE_SIGNS, bc, nx, ny, 2)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def _fill_gc1_scalar(mem, signs, bc, nx, ny):
    """Independent scalar statement of the one-ghost-cell fill: read index
    wrapped for periodic sides, then N, S, E, W mirrors/copies."""
    n, s, e, w = O.bc_north(bc), O.bc_south(bc), O.bc_east(bc), O.bc_west(bc)
    out = []
    for a, (sx, sy) in zip(mem, signs):
        q = np.empty_like(a)
        for j in range(ny + 2):
            for i in range(nx + 2):
                jj, ii = j, i
                if j == 0 and s == P:
                    jj = ny
                if j == ny + 1 and n == P:
                    jj = 1
                if i == 0 and w == P:
                    ii = nx
                if i == nx + 1 and e == P:
                    ii = 1
                q[j, i] = a[jj, ii]
        for i in range(nx + 2):
            if n == R:
                q[ny + 1, i] = sy * q[ny, i]
            if n == F:
                q[ny + 1, i] = q[ny, i]
        for i in range(nx + 2):
            if s == R:
                q[0, i] = sy * q[1, i]
            if s == F:
                q[0, i] = q[1, i]
        for j in range(ny + 2):
            if e == R:
                q[j, nx + 1] = sx * q[j, nx]
            if e == F:
                q[j, nx + 1] = q[j, nx]
        for j in range(ny + 2):
            if w == R:
                q[j, 0] = sx * q[j, 1]
            if w == F:
                q[j, 0] = q[j, 1]
        out.append(q)
    return out


@pytest.mark.parametrize("bc", [O.pack_bc(R, R, R, R), O.pack_bc(P, P, P, P), O.pack_bc(R, P, D, F),
                                O.pack_bc(D, D, D, D), O.pack_bc(P, R, F, P)])
def test_ghost_fill_gc1_equals_scalar_statement(bc):
    nx, ny = 6, 5
    rng = np.random.default_rng(2)
    mem = [rng.standard_normal((ny + 2, nx + 2)) for _ in range(3)]
    a = U.fill_ghosts_gc(mem, O.SWE_SIGNS, bc, nx, ny, 1)
    b = _fill_gc1_scalar(mem, O.SWE_SIGNS, bc, nx, ny)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


# --------------------------------------------------------------------------
# fluxes
# --------------------------------------------------------------------------

@pytest.mark.parametrize("flux", ["lxf1", "lxf2", "lxw2", "force"])
def test_centred_flux_consistency(flux):
    """F(Q, Q) = F(Q) exactly: every centred flux reduces to the physical flux."""
    q = tuple(np.asarray(v, dtype=np.float32) for v in ([1.3, 0.7], [0.4, -0.2], [-0.1, 0.3]))
    g, dx, dt = np.float32(9.81), np.float32(0.5), np.float32(0.01)
    fn = {"lxf1": U.lxf_1d_flux, "lxf2": U.lxf_2d_flux, "lxw2": U.lxw2_1d_flux, "force": U.force_1d_flux}[flux]
    got = fn(q, q, g, dx, dt)
    want = O.swe_F(*q, g)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_centred_fluxes_closed_form():
    """Two fixed states, the Toro (2001) formulas evaluated by hand in float64."""
    ql = (np.float64(2.0), np.float64(1.0), np.float64(0.5))
    qr = (np.float64(1.0), np.float64(-0.5), np.float64(0.25))
    g, dx, dt = 10.0, 0.2, 0.01
    Fl = (1.0, 1.0 / 2.0 + 0.5 * 10 * 4.0, 0.5 / 2.0)
    Fr = (-0.5, 0.25 / 1.0 + 0.5 * 10 * 1.0, -0.125 / 1.0)
    lxf1 = [0.5 * (a + b) + (dx / (2 * dt)) * (c - d) for a, b, c, d in zip(Fl, Fr, ql, qr)]
    lxf2 = [0.5 * (a + b) + (dx / (4 * dt)) * (c - d) for a, b, c, d in zip(Fl, Fr, ql, qr)]
    qs = [0.5 * (c + d) + (dt / (2 * dx)) * (a - b) for a, b, c, d in zip(Fl, Fr, ql, qr)]
    lw = (qs[1], qs[1] ** 2 / qs[0] + 0.5 * g * qs[0] ** 2, qs[1] * qs[2] / qs[0])
    assert np.allclose(U.lxf_1d_flux(ql, qr, g, dx, dt), lxf1, rtol=1e-14)
    assert np.allclose(U.lxf_2d_flux(ql, qr, g, dx, dt), lxf2, rtol=1e-14)
    assert np.allclose(U.lxw2_1d_flux(ql, qr, g, dx, dt), lw, rtol=1e-14)
    assert np.allclose(U.force_1d_flux(ql, qr, g, dx, dt), [0.5 * (a + b) for a, b in zip(lxf1, lw)], rtol=1e-14)


# --------------------------------------------------------------------------
# whole-step invariants
# --------------------------------------------------------------------------

@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("bc", [O.pack_bc(R, R, R, R), O.pack_bc(P, P, P, P), O.pack_bc(F, F, R, P)])
def test_lake_at_rest(scheme, bc):
    nx, ny = 11, 9
    h = np.full((ny, nx), 1.7, dtype=np.float32)
    z = np.zeros_like(h)
    sim = _run(scheme, [h, z, z], bc, 5, 0.05, dtype=np.float32)
    out = sim.download()
    assert np.array_equal(out[0], h) and not out[1].any() and not out[2].any()


@pytest.mark.parametrize("scheme", SCHEMES)
def test_mass_conservation_periodic(scheme):
    nx, ny = 13, 10
    f = _state(nx, ny, 3)
    sim = _run(scheme, f, O.pack_bc(P, P, P, P), 6, 0.02)
    out = sim.download()
    assert abs(out[0].sum() - f[0].sum()) < 1e-11 * f[0].sum()
    assert abs(out[1].sum() - f[1].sum()) < 1e-10 and abs(out[2].sum() - f[2].sum()) < 1e-10


@pytest.mark.parametrize("scheme", SCHEMES)
def test_mirror_symmetry_x(scheme):
    """Mirroring the initial state in x (hu negated) mirrors the solution."""
    nx, ny = 12, 9
    f = _state(nx, ny, 4)
    bc = O.pack_bc(R, R, R, R)
    a = _run(scheme, f, bc, 5, 0.02).download()
    m = [f[0][:, ::-1].copy(), -f[1][:, ::-1], f[2][:, ::-1].copy()]
    b = _run(scheme, m, bc, 5, 0.02).download()
    assert np.allclose(a[0], b[0][:, ::-1], rtol=0, atol=1e-12)
    assert np.allclose(a[1], -b[1][:, ::-1], rtol=0, atol=1e-12)
    assert np.allclose(a[2], b[2][:, ::-1], rtol=0, atol=1e-12)


@pytest.mark.parametrize("scheme", ["lxf", "kp07"])
def test_unsplit_transposition(scheme):
    """The unsplit schemes treat x and y alike: the transposed state (hu, hv
    swapped) evolves into the transposed solution."""
    nx, ny = 10, 10
    f = _state(nx, ny, 5)
    bc = O.pack_bc(R, R, P, P)
    a = _run(scheme, f, bc, 4, 0.02).download()
    t = [f[0].T.copy(), f[2].T.copy(), f[1].T.copy()]
    b = _run(scheme, t, O.pack_bc(P, P, R, R), 4, 0.02).download()
    assert np.allclose(a[0], b[0].T, rtol=0, atol=1e-12)
    assert np.allclose(a[1], b[2].T, rtol=0, atol=1e-12)
    assert np.allclose(a[2], b[1].T, rtol=0, atol=1e-12)


# --------------------------------------------------------------------------
# 1-D reductions against scalar loops
# --------------------------------------------------------------------------

def _F1(h, hu, g):
    return (hu, hu * hu / h + 0.5 * g * h * h, 0.0)


def _scalar_1d(scheme, h, hu, g, dx, dt, theta=1.3, order=2):
    """Independent scalar statement of one step on a 1-D row with reflective
    ends (ghosts mirror, hu negated); the 2-D schemes reduce to it on a
    y-uniform state with zero hv, except for LxF's 2-D diffusion dx/(4dt)."""
    n = len(h)

    def ghosted(h, hu, gc):
        H = [0.0] * (n + 2 * gc)
        HU = [0.0] * (n + 2 * gc)
        for i in range(n):
            H[i + gc], HU[i + gc] = h[i], hu[i]
        for k in range(gc):
            H[gc - 1 - k], HU[gc - 1 - k] = H[gc + k], -HU[gc + k]
            H[n + gc + k], HU[n + gc + k] = H[n + gc - 1 - k], -HU[n + gc - 1 - k]
        return H, HU

    def phys(a, b):
        f = _F1(a, b, g)
        return f[0], f[1]

    if scheme in ("lxf", "force", "hll"):
        H, HU = ghosted(h, hu, 1)
        flux = []
        for i in range(n + 1):
            hl, ql, hr, qr = H[i], HU[i], H[i + 1], HU[i + 1]
            fl, fr = phys(hl, ql), phys(hr, qr)
            if scheme == "lxf":
                k = dx / (4 * dt)
                flux.append(tuple(0.5 * (fl[c] + fr[c]) + k * ((hl, ql)[c] - (hr, qr)[c]) for c in range(2)))
            elif scheme == "force":
                k1, k2 = dx / (2 * dt), dt / (2 * dx)
                lf = [0.5 * (fl[c] + fr[c]) + k1 * ((hl, ql)[c] - (hr, qr)[c]) for c in range(2)]
                hs = 0.5 * (hl + hr) + k2 * (fl[0] - fr[0])
                qs = 0.5 * (ql + qr) + k2 * (fl[1] - fr[1])
                lw = phys(hs, qs)
                flux.append(tuple(0.5 * (lf[c] + lw[c]) for c in range(2)))
            else:
                ul, ur = ql / hl, qr / hr
                cl, cr = np.sqrt(g * hl), np.sqrt(g * hr)
                hd = 0.5 * (hl + hr) - 0.25 * (ur - ul) * (hl + hr) / (cl + cr)
                sl = ul - cl * (np.sqrt(0.5 * (hd + hl) * hd / hl ** 2) if hd > hl else 1.0)
                sr = ur + cr * (np.sqrt(0.5 * (hd + hr) * hd / hr ** 2) if hd > hr else 1.0)
                if sl >= 0:
                    flux.append(fl)
                elif sr <= 0:
                    flux.append(fr)
                else:
                    flux.append(tuple((sr * fl[c] - sl * fr[c] + sr * sl * ((hr, qr)[c] - (hl, ql)[c])) / (sr - sl)
                                      for c in range(2)))
        return ([h[i] + (flux[i][0] - flux[i + 1][0]) * dt / dx for i in range(n)],
                [hu[i] + (flux[i][1] - flux[i + 1][1]) * dt / dx for i in range(n)])

    def kp07_stage(h, hu):
        H, HU = ghosted(h, hu, 2)

        def mm(l, c, r):
            b, m, f = theta * (c - l), 0.5 * (r - l), theta * (r - c)
            if b > 0 and m > 0 and f > 0:
                return min(b, m, f)
            if b < 0 and m < 0 and f < 0:
                return max(b, m, f)
            return 0.0
        S = {i: (mm(H[i - 1], H[i], H[i + 1]), mm(HU[i - 1], HU[i], HU[i + 1])) for i in range(1, n + 3)}
        flux = []
        for i in range(1, n + 2):  # interface (i, i+1) of the padded row
            qm = (H[i] + 0.5 * S[i][0], HU[i] + 0.5 * S[i][1])
            qp = (H[i + 1] - 0.5 * S[i + 1][0], HU[i + 1] - 0.5 * S[i + 1][1])
            fm, fp = phys(*qm), phys(*qp)
            um, up = qm[1] / qm[0], qp[1] / qp[0]
            cm, cp = np.sqrt(g * qm[0]), np.sqrt(g * qp[0])
            am = min(um - cm, up - cp, 0.0)
            ap = max(um + cm, up + cp, 0.0)
            flux.append(tuple(((ap * fm[c] - am * fp[c]) + ap * am * (qp[c] - qm[c])) / (ap - am) for c in range(2)))
        return ([h[i] + (flux[i][0] - flux[i + 1][0]) * dt / dx for i in range(n)],
                [hu[i] + (flux[i][1] - flux[i + 1][1]) * dt / dx for i in range(n)])

    h1, q1 = kp07_stage(list(h), list(hu))
    if order == 1:
        return h1, q1
    h2, q2 = kp07_stage(h1, q1)
    return [0.5 * (a + b) for a, b in zip(h, h2)], [0.5 * (a + b) for a, b in zip(hu, q2)]


@pytest.mark.parametrize("scheme", SCHEMES)
def test_one_dimensional_reduction_equals_scalar_loop(scheme):
    nx, ny = 17, 3
    x = (np.arange(nx) + 0.5) / nx
    h1 = 1.0 + 0.4 * np.exp(-((x - 0.4) / 0.15) ** 2)
    hu1 = 0.3 * np.sin(2 * np.pi * x)
    h = np.repeat(h1[None, :], ny, 0)
    hu = np.repeat(hu1[None, :], ny, 0)
    dt, dx = 0.01, 1.0 / nx
    sim = _run(scheme, [h, hu, np.zeros_like(h)], O.pack_bc(P, P, R, R), 1, dt, dx=dx, dy=dx)
    out = sim.download()
    want_h, want_hu = _scalar_1d(scheme, list(h1), list(hu1), 9.81, dx, dt)
    for j in range(ny):
        assert np.allclose(out[0][j], want_h, rtol=0, atol=1e-13), scheme
        assert np.allclose(out[1][j], want_hu, rtol=0, atol=1e-13), scheme
    assert np.array_equal(out[2], np.zeros_like(h))


def test_kp07_order1_is_one_stage():
    nx, ny = 9, 8
    f = _state(nx, ny, 6)
    bc = O.pack_bc(R, P, R, P)
    one = U.OracleUnsplitSimulator("kp07", f, nx, ny, 1.0, 1.0, 9.81, order=1, bc=bc, dtype=np.float64)
    one.step(0.01)
    U0 = [U.pad_gc(a, np.float64, 2) for a in f]
    U1 = [np.zeros_like(a) for a in U0]
    U.kp07_substep(U0, U1, 9.81, 1.3, 1.0, 1.0, 0.01, bc, 0, 1)
    for a, b in zip(one.download(), U1):
        assert np.array_equal(a, b[2:-2, 2:-2])
    assert one.compute_dt() == one.cfl  # order 1: factor 0.5**0


def test_cfl_is_min_over_written_cells():
    nx, ny = 8, 6
    f = _state(nx, ny, 7, np.float32)
    for scheme in SCHEMES:
        sim = _run(scheme, f, O.pack_bc(R, R, R, R), 1, 0.01, dtype=np.float32)
        out = sim.download()
        if scheme == "kp07":
            continue  # the minimum is of the pre-average state, checked on the GPU against the oracle
        assert sim.cfl == np.min(O.swe_cfl(out[0], out[1], out[2], np.float32(1.0), np.float32(1.0),
                                           np.float32(9.81)))


# --------------------------------------------------------------------------
# accuracy
# --------------------------------------------------------------------------

def _stoker(scheme, nx):
    ny = 3
    x = (np.arange(nx) + 0.5) * (10.0 / nx)
    h = np.where(x < 5.0, 0.005, 0.001)[None, :].repeat(ny, 0)
    z = np.zeros_like(h)
    dx = 10.0 / nx
    sim = U.OracleUnsplitSimulator(scheme, [h, z, z], nx, ny, dx, dx, 9.81, bc=O.pack_bc(P, P, R, R),
                                   dtype=np.float64)
    t = 0.0
    while t < 6.0:
        dt = min(sim.compute_dt() * 0.9, 6.0 - t)
        sim.step(dt)
        t += dt
    return sim.download()[0][0]


@pytest.mark.parametrize("scheme", SCHEMES)
def test_stoker_dam_break_against_swashes(scheme):
    from tests.test_oracle_pins import _swashes
    errs = []
    for nx in (128, 256):
        xs, h_ref, _ = _swashes(nx)
        errs.append(np.mean(np.abs(_stoker(scheme, nx) - h_ref)) / np.mean(h_ref))
    assert errs[0] > errs[1], errs
    assert errs[1] < (0.03 if scheme == "kp07" else 0.08), errs


def _smooth(scheme, n, t_end=0.3):
    L = 1.0
    x = (np.arange(n) + 0.5) * (L / n)
    X, Y = np.meshgrid(x, x)
    h = 1.0 + 0.05 * np.exp(-((X - 0.5) ** 2 + (Y - 0.5) ** 2) / 0.02)
    z = np.zeros_like(h)
    sim = U.OracleUnsplitSimulator(scheme, [h, z, z], n, n, L / n, L / n, 1.0, bc=O.pack_bc(P, P, P, P),
                                   dtype=np.float64)
    dt = 0.2 * (L / n) / 1.05
    steps = int(np.ceil(t_end / dt))
    for _ in range(steps):
        sim.step(t_end / steps)
    return sim.download()[0]


@pytest.mark.parametrize("scheme,lo,hi", [("kp07", 1.6, 3.0), ("lxf", 0.6, 1.4), ("hll", 0.6, 1.4),
                                          ("force", 0.6, 1.4)])
def test_observed_order(scheme, lo, hi):
    ref = _smooth(scheme, 128)
    errs = []
    for n in (16, 32, 64):
        c = _smooth(scheme, n)
        f = 128 // n
        errs.append(np.mean(np.abs(c - ref.reshape(n, f, n, f).mean(axis=(1, 3)))))
    order = np.log2(errs[1] / errs[2])
    assert lo < order < hi, (errs, order)
