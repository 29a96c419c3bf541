"""Initial conditions of the reference's test cases, with the API of
GPUSimulators/helpers/InitialConditions.py (getExtent, downsample, bump,
genShockBubble, genKelvinHelmholtz, genRayleighTaylor).

Host-side NumPy setup code: each generator returns the arrays (and, for the
Euler cases, the keyword arguments of EE2D_KP07_dimsplit) that the reference's
generator returns for the same arguments.  ``grid`` is an MPIGrid (or any
object with ``grid`` = (gx, gy) and ``getCoordinate(index)``) when the domain
is split over ranks.
"""

import numpy as np

from ..Simulator import BoundaryCondition


def getExtent(width, height, nx, ny, grid, index=None):
    """[x0, x1, y0, y1, dx, dy]: first and last cell centres and the cell size of
    this rank's part of a width x height domain (InitialConditions.py:28-54).
    The last y centre uses dx, as the reference does."""
    if grid is None:
        dx, dy = width / nx, height / ny
        return [0.5 * dx, width - 0.5 * dx, 0.5 * dy, height - 0.5 * dy, dx, dy]
    gx, gy = grid.grid[0], grid.grid[1]
    i, j = grid.getCoordinate(index) if index is not None else grid.getCoordinate()
    dx = (width / gx) / nx
    dy = (height / gy) / ny
    return [width * i / gx + 0.5 * dx, width * (i + 1) / gx - 0.5 * dx,
            height * j / gy + 0.5 * dy, height * (j + 1) / gy - 0.5 * dx, dx, dy]


def downsample(highres_solution, x_factor, y_factor=None):
    """Block means of a (ny, nx) array over x_factor x y_factor cells
    (InitialConditions.py:57-74); 1-D input is treated as one row."""
    if y_factor is None:
        y_factor = x_factor
    a = np.asarray(highres_solution)
    if a.ndim == 1:
        a = a.reshape((1, a.size))
    xf, yf = int(round(x_factor)), int(round(y_factor))
    assert a.shape[1] % xf == 0 and a.shape[0] % yf == 0
    if xf * yf == 1:
        return a
    ny, nx = a.shape[0] // yf, a.shape[1] // xf
    return a.reshape(ny, yf, nx, xf).mean(3).mean(1)


def bump(nx, ny, width, height, bump_size=None, ref_nx=None, ref_ny=None, x_center=0.5, y_center=0.5,
         h_ref=0.5, h_amp=0.1, u_ref=0.0, u_amp=0.1, v_ref=0.0, v_amp=0.1):
    """Cosine bump of radius ``bump_size`` (default width/5) in h, u and v on a
    ``ref_nx`` x ``ref_ny`` grid, averaged down to nx x ny
    (InitialConditions.py:79-132).  Returns h, hu = h*u, hv = h*v, dx, dy."""
    ref_nx = nx if ref_nx is None else ref_nx
    ref_ny = ny if ref_ny is None else ref_ny
    assert ref_nx >= nx and ref_ny >= ny
    if bump_size is None:
        bump_size = width / 5.0
    rdx, rdy = width / float(ref_nx), height / float(ref_ny)
    xc, yc = rdx * ref_nx * x_center, rdy * ref_ny * y_center
    x = rdx * (np.arange(0, ref_nx, dtype=np.float32) + 0.5) - xc
    y = rdy * (np.arange(0, ref_ny, dtype=np.float32) + 0.5) - yc
    xv, yv = np.meshgrid(x, y, sparse=False, indexing="xy")
    r = np.sqrt(xv ** 2 + yv ** 2)
    shape = 0.5 * (1.0 + np.cos(np.pi * r / bump_size)) * (r < bump_size)
    fx, fy = ref_nx / nx, ref_ny / ny
    h = downsample(h_ref + h_amp * shape, fx, fy)
    hu = downsample(u_ref + u_amp * shape, fx, fy) * h
    hv = downsample(v_ref + v_amp * shape, fx, fy) * h
    return h, hu, hv, width / nx, height / ny


def _cell_centres(x0, x1, n):
    return np.linspace(x0, x1, n, dtype=np.float32)


def genShockBubble(nx, ny, gamma, grid=None):
    """Shock-bubble interaction on a 4 x 1 domain (InitialConditions.py:135-195):
    a light bubble (rho 0.1, radius 0.25 at (0.5, 0.5)) in still gas (rho 1,
    p 1), hit by the shocked state rho 3.8125, u 2.57669, p 10 entering from
    x < 0.1.  Reflective north/south, periodic east/west, no gravity."""
    width, height, g = 4.0, 1.0, 0.0
    x0, x1, y0, y1, dx, dy = getExtent(width, height, nx, ny, grid)
    xv, yv = np.meshgrid(_cell_centres(x0, x1, nx), _cell_centres(y0, y1, ny), sparse=False, indexing="xy")
    rho = np.ones((ny, nx), dtype=np.float32)
    u = np.zeros((ny, nx), dtype=np.float32)
    v = np.zeros((ny, nx), dtype=np.float32)
    p = np.ones((ny, nx), dtype=np.float32)
    rho = np.where(np.sqrt((xv - 0.5) ** 2 + (yv - 0.5) ** 2) <= 0.25, 0.1, rho)
    left = xv < 0.1
    rho = np.where(left, 3.81250, rho)
    u = np.where(left, 2.57669, u)
    p = np.where(left, 10.0, p)
    E = 0.5 * rho * (u ** 2 + v ** 2) + p / (gamma - 1.0)
    T = BoundaryCondition.Type
    bc = BoundaryCondition({"north": T.Reflective, "south": T.Reflective, "east": T.Periodic, "west": T.Periodic})
    return {"rho": rho, "rho_u": rho * u, "rho_v": rho * v, "E": E, "nx": nx, "ny": ny, "dx": dx, "dy": dy,
            "g": g, "gamma": gamma, "boundary_conditions": bc}


def genKelvinHelmholtz(nx, ny, gamma, roughness=0.125, grid=None, index=None):
    """Kelvin-Helmholtz shear layers on a 2 x 1 periodic domain
    (InitialConditions.py:198-300): a dense band (rho 2, u -0.5) between
    y = 0.25 and y = 0.75, perturbed by smooth random curves of amplitude one
    cell (``roughness`` x nx control points, interpolated), in light gas
    (rho 1, u 0.5); p = 2.5.  Draws from NumPy's global random state, as the
    reference does; gamma is fixed at 1.4 as there."""
    def smooth_random(n_pts, n):
        n = max(1, min(n, n_pts))
        if n == n_pts:
            return np.random.random(n_pts) - 0.5
        from scipy.interpolate import interp1d
        kind = {1: "nearest", 2: "linear", 3: "quadratic"}.get(n, "cubic")
        f = interp1d(np.linspace(0.0, 1.0, n), np.random.random(n) - 0.5, kind=kind)
        return f(np.linspace(0.0, 1.0, n_pts))

    def zones(n):
        x0, x1, y0, y1, _, dy = getExtent(1.0, 1.0, nx, ny, grid, index)
        _, y = np.meshgrid(np.linspace(x0, x1, nx), np.linspace(y0, y1, ny))
        z = np.zeros((ny, nx), dtype=np.int32)
        z = np.where(y > 0.25 + smooth_random(nx, n) * dy, z, 1)
        z = np.where(y < 0.75 + smooth_random(nx, n) * dy, z, 1)
        return z

    gamma = 1.4
    z = zones(max(1, min(nx, int(nx * roughness))))
    rho = np.where(z == 0, 1.0, 2.0).astype(np.float32)
    u = np.where(z == 0, 0.5, -0.5).astype(np.float32)
    v = np.zeros((ny, nx), dtype=np.float32)
    p = 2.5 * np.ones((ny, nx), dtype=np.float32)
    E = 0.5 * rho * (u ** 2 + v ** 2) + p / (gamma - 1.0)
    _, _, _, _, dx, dy = getExtent(2.0, 1.0, nx, ny, grid, index)
    T = BoundaryCondition.Type
    bc = BoundaryCondition({"north": T.Periodic, "south": T.Periodic, "east": T.Periodic, "west": T.Periodic})
    return {"rho": rho, "rho_u": rho * u, "rho_v": rho * v, "E": E, "nx": nx, "ny": ny, "dx": dx, "dy": dy,
            "g": 0.0, "gamma": gamma, "boundary_conditions": bc}


def genRayleighTaylor(nx, ny, gamma, version=0, grid=None):
    """Rayleigh-Taylor column on 0.5 x 1.5 with g = 0.1
    (InitialConditions.py:304-352): rho 2 above, 1 below an interface at
    y = 0.01 cos(4 pi |x|) (version 0) or flat with v = 0.01 (1 + cos(4 pi
    x)) / 4 (version 1); hydrostatic p = 2.5 - rho g y; reflective walls."""
    width, height, g = 0.5, 1.5, 0.1
    x0, x1, y0, y1, dx, dy = getExtent(width, height, nx, ny, grid)
    x = _cell_centres(x0, x1, nx) - width * 0.5
    y = _cell_centres(y0, y1, ny) - height * 0.5
    xv, yv = np.meshgrid(x, y, sparse=False, indexing="xy")
    u = np.zeros((ny, nx), dtype=np.float32)
    v = np.zeros((ny, nx), dtype=np.float32)
    if version == 0:
        threshold = 0.01 * np.cos(2 * np.pi * np.abs(x) / 0.5)
        rho = np.where(yv <= threshold, 1.0, 2.0).astype(np.float32)
    elif version == 1:
        rho = np.where(yv <= 0.0, 1.0, 2.0).astype(np.float32)
        v = 0.01 * (1.0 + np.cos(2 * np.pi * xv / 0.5)) / 4
    else:
        raise AssertionError("Invalid version")
    p = 2.5 - rho * g * yv
    E = 0.5 * rho * (u ** 2 + v ** 2) + p / (gamma - 1.0)
    T = BoundaryCondition.Type
    bc = BoundaryCondition({"north": T.Reflective, "south": T.Reflective, "east": T.Reflective,
                            "west": T.Reflective})
    return {"rho": rho, "rho_u": rho * u, "rho_v": rho * v, "E": E, "nx": nx, "ny": ny, "dx": dx, "dy": dy,
            "g": g, "gamma": gamma, "boundary_conditions": bc}
