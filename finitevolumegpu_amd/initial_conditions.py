"""Seeded initial conditions for the benchmark and test workloads.

Every generator draws from ``numpy.random.default_rng(seed)`` and returns
float32 C-contiguous ``(ny, nx)`` interior fields (the reference's upload
format, GPUSimulators/Common.py:523-531).  Cell centres are at ``(i+0.5)*dx``.

* ``circular_dam_break``: the 256x256 shallow-water dam break (h=1 background,
  +0.5 inside a disc, uniform noise on h, zero momenta).
* ``dam_break_bumps``: the same disc plus Gaussian bumps (the 4096x4096
  single-grid workload and its 512x512 reduced copy).
* ``rayleigh_taylor``: the Euler-with-gravity Rayleigh-Taylor setup in the
  shape of GPUSimulators/helpers/InitialConditions.py:304-355, with a seeded
  multi-mode interface.
* ``ensemble_instances`` / ``serving_requests``: the batched-ensemble and
  serving workloads.
"""

import math

import numpy as np


def _centres(n, d=1.0):
    return (np.arange(n, dtype=np.float64) + 0.5) * d


def _disk(n, cx, cy, r, inside, outside=1.0):
    """n x n float32 field: ``inside`` at the cell centres with
    (x-cx)^2 < r^2 - (y-cy)^2, ``outside`` elsewhere.  Each row's disk cells are
    one contiguous run [lo, hi), found by binary search on the two monotone
    halves of (x-cx)^2, so the field is written in whole-array passes (no
    per-row Python loop holding the interpreter lock)."""
    x = _centres(n)
    dx2 = (x - cx) ** 2
    thr = r * r - (x - cy) ** 2
    i0 = int(np.argmin(dx2))
    left = dx2[:i0 + 1]   # non-increasing
    right = dx2[i0:]      # non-decreasing
    lo = np.searchsorted(-left, -thr, side="right")       # count of left cells with dx2 >= thr
    hi = i0 + np.searchsorted(right, thr, side="left")    # i0 + count of right cells with dx2 < thr
    h = np.full((n, n), outside, dtype=np.float32)
    rows = np.flatnonzero(hi > lo)
    if rows.size:
        j0, j1 = rows[0], rows[-1] + 1  # the disk's bounding box
        c0, c1 = int(lo[j0:j1].min()), int(hi[j0:j1].max())
        cols = np.arange(c0, c1)
        mask = (cols[None, :] >= lo[j0:j1, None]) & (cols[None, :] < hi[j0:j1, None])
        box = h[j0:j1, c0:c1]
        box[mask] = np.float32(inside)
    return h


def circular_dam_break(nx=256, ny=256, radius=32.0, depth=1.0, jump=0.5,
                       noise=1e-3, seed=0, centre=None):
    """h = depth + jump*[r < radius] + U(-noise, noise); hu = hv = 0 (dx = dy = 1)."""
    rng = np.random.default_rng(seed)
    cx, cy = (nx / 2.0, ny / 2.0) if centre is None else centre
    x = _centres(nx)
    y = _centres(ny)
    r = np.hypot(x[None, :] - cx, y[:, None] - cy)
    h = depth + jump * (r < radius)
    if noise:
        h = h + rng.uniform(-noise, noise, size=(ny, nx))
    h = np.ascontiguousarray(h, dtype=np.float32)
    z = np.zeros((ny, nx), dtype=np.float32)
    return h, z, z.copy()


def dam_break_bumps(n=4096, seed=0, nbumps=16, noise=1e-3):
    """Dam break scaled to an n x n grid (radius n/8, centre n/2) plus ``nbumps``
    Gaussian bumps a*exp(-r^2/(2w^2)), a~U(0.05,0.3), w~U(20,120)*n/4096, centres
    uniform over the domain."""
    rng = np.random.default_rng(seed)
    scale = n / 4096.0
    x = _centres(n)
    y = _centres(n)
    r = np.hypot(x[None, :] - n / 2.0, y[:, None] - n / 2.0)
    h = 1.0 + 0.5 * (r < n / 8.0)
    h = h + rng.uniform(-noise, noise, size=(n, n))
    amp = rng.uniform(0.05, 0.3, size=nbumps)
    wid = rng.uniform(20.0, 120.0, size=nbumps) * scale
    cx = rng.uniform(0.0, n, size=nbumps)
    cy = rng.uniform(0.0, n, size=nbumps)
    for a, w, px, py in zip(amp, wid, cx, cy):
        gx = np.exp(-((x - px) ** 2) / (2.0 * w * w))
        gy = np.exp(-((y - py) ** 2) / (2.0 * w * w))
        h += a * gy[:, None] * gx[None, :]
    h = np.ascontiguousarray(h, dtype=np.float32)
    z = np.zeros((n, n), dtype=np.float32)
    return h, z, z.copy()


def rayleigh_taylor(nx, ny, gamma=1.4, g=0.1, width=0.5, seed=0, amplitude=0.01, modes=4,
                    p_interface=2.5):
    """Heavy (rho=2) fluid above light (rho=1) fluid, hydrostatic pressure from
    ``p_interface`` at the interface y_i(x) = Ly/2 + sum_m a_m cos(2 pi m x/Lx + phi_m),
    sum |a_m| = amplitude*Ly, zero velocity.  dx = dy = width/nx.

    Returns (rho, rho_u, rho_v, E, dx, dy)."""
    rng = np.random.default_rng(seed)
    dx = width / nx
    dy = dx
    Lx, Ly = nx * dx, ny * dy
    w = rng.uniform(0.2, 1.0, size=modes)
    a = amplitude * Ly * w / w.sum()
    phi = rng.uniform(0.0, 2.0 * math.pi, size=modes)
    x = _centres(nx, dx)
    y = _centres(ny, dy)
    yi = Ly / 2.0 + sum(a[m] * np.cos(2.0 * math.pi * (m + 1) * x / Lx + phi[m]) for m in range(modes))
    above = y[:, None] > yi[None, :]
    rho = np.where(above, 2.0, 1.0)
    p = p_interface - rho * g * (y[:, None] - yi[None, :])
    E = p / (gamma - 1.0)
    f32 = lambda v: np.ascontiguousarray(v, dtype=np.float32)
    z = np.zeros((ny, nx), dtype=np.float32)
    return f32(rho), z, z.copy(), f32(E), dx, dy


def ensemble_instances(count=512, seed=0, side_min=64, side_max=1024,
                       t_min=5.0, t_max=50.0):
    """Batched-ensemble workload: ``count`` shallow-water dam breaks with
    side log-uniform in [side_min, side_max], radius U(0.05,0.3)*side, depth ratio
    U(1.1,3.0), random centre, reflective or outflow walls (50/50), t_end U(t_min,t_max).

    Returns a list of dicts with keys nx, ny, h, hu, hv, t_end, bc ('reflective'|'outflow')."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        s = int(round(math.exp(rng.uniform(math.log(side_min), math.log(side_max)))))
        frac = rng.uniform(0.05, 0.3)
        ratio = rng.uniform(1.1, 3.0)
        r = frac * s
        cx = rng.uniform(r, s - r)
        cy = rng.uniform(r, s - r)
        bc = "reflective" if rng.uniform() < 0.5 else "outflow"
        t_end = rng.uniform(t_min, t_max)
        h = _disk(s, cx, cy, r, ratio)
        z = np.zeros((s, s), dtype=np.float32)
        out.append({"nx": s, "ny": s, "h": h, "hu": z, "hv": z.copy(), "t_end": float(t_end), "bc": bc})
    return out


def serving_requests(count=2000, seed=0, rate=None, side_min=128, side_max=2048,
                     t_min=2.0, t_max=40.0):
    """Serving workload: Poisson arrivals (``rate`` requests/s; ``None`` = all at t=0),
    60 % KP07_dimsplit, 20 % HLL2 or WAF, 20 % EE2D_KP07_dimsplit, side log-uniform in
    [side_min, side_max].  Initial data are generated lazily by ``materialise``.

    Returns a list of dicts: id, arrival, scheme, n (= nx = ny), t_end, seed."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(count):
        if rate:
            t += rng.exponential(1.0 / rate)
        u = rng.uniform()
        if u < 0.6:
            scheme = "kp07_dimsplit"
        elif u < 0.8:
            scheme = "hll2" if rng.uniform() < 0.5 else "waf"
        else:
            scheme = "ee2d_kp07_dimsplit"
        n = int(round(math.exp(rng.uniform(math.log(side_min), math.log(side_max)))))
        reqs.append({"id": i, "arrival": t, "scheme": scheme, "n": n, "nx": n, "ny": n,
                     "t_end": float(rng.uniform(t_min, t_max)), "seed": int(rng.integers(0, 2**31 - 1))})
    return reqs


def materialise(req):
    """Initial fields for one serving request.  SWE: a dam break (as the ensemble);
    Euler: a Rayleigh-Taylor column on the unit-width domain whose t_end is scaled
    by dx so the step count is comparable with the SWE requests."""
    rng = np.random.default_rng(req["seed"])
    n = req["n"]
    if req["scheme"] == "ee2d_kp07_dimsplit":
        rho, ru, rv, E, dx, dy = rayleigh_taylor(n, n, width=1.0, seed=req["seed"])
        return {"fields": [rho, ru, rv, E], "dx": dx, "dy": dy, "g": 0.1, "gamma": 1.4,
                "t_end": req["t_end"] * dx, "bc": "ns_reflective_ew_periodic"}
    frac = rng.uniform(0.05, 0.3)
    ratio = rng.uniform(1.1, 3.0)
    r = frac * n
    cx = rng.uniform(r, n - r)
    cy = rng.uniform(r, n - r)
    h = _disk(n, cx, cy, r, ratio)
    z = np.zeros((n, n), dtype=np.float32)
    return {"fields": [h, z, z.copy()], "dx": 1.0, "dy": 1.0, "g": 9.81, "gamma": None,
            "t_end": req["t_end"], "bc": "reflective"}
