"""The reference's initial conditions (finitevolumegpu_amd/helpers/
InitialConditions.py, API of GPUSimulators/helpers/InitialConditions.py):
extents, block means, the cosine bump's integral, and the physical states of
the shock-bubble, Kelvin-Helmholtz and Rayleigh-Taylor cases."""

import numpy as np
import pytest

from finitevolumegpu_amd.helpers import InitialConditions as I


class _Grid:
    def __init__(self, gx, gy, rank):
        self.grid = np.array([gx, gy])
        self.rank = rank

    def getCoordinate(self, rank=None):
        rank = self.rank if rank is None else rank
        return rank % self.grid[0], rank // self.grid[0]


def test_get_extent_single_and_split():
    x0, x1, y0, y1, dx, dy = I.getExtent(4.0, 1.0, 40, 10, None)
    assert (dx, dy) == (0.1, 0.1) and np.isclose(x0, 0.05) and np.isclose(x1, 3.95) and np.isclose(y1, 0.95)
    # two ranks in x: the halves tile the domain
    e0 = I.getExtent(4.0, 1.0, 20, 10, _Grid(2, 1, 0))
    e1 = I.getExtent(4.0, 1.0, 20, 10, _Grid(2, 1, 1))
    assert np.isclose(e0[4], 0.1) and np.isclose(e0[1] + e0[4], e1[0])
    assert np.isclose(e1[1], 3.95)


def test_downsample_block_means():
    a = np.arange(24, dtype=np.float64).reshape(4, 6)
    d = I.downsample(a, 3, 2)
    assert d.shape == (2, 2)
    assert d[0, 0] == np.mean(a[0:2, 0:3]) and d[1, 1] == np.mean(a[2:4, 3:6])
    assert np.isclose(d.mean(), a.mean())
    assert I.downsample(a, 1) is a or np.array_equal(I.downsample(a, 1), a)
    assert I.downsample(np.arange(6.0), 2, 1).shape == (1, 3)  # a row: y_factor 1


def test_bump_integral_and_refinement():
    """The cosine bump 0.5 (1 + cos(pi r / R)) integrates to pi R^2/2 - 2 R^2/pi;
    a bump generated on a finer grid and averaged down equals the downsample of
    the fine one."""
    R, W = 20.0, 100.0
    h, hu, hv, dx, dy = I.bump(512, 512, W, W, bump_size=R)
    mass = (h - 0.5).sum() * dx * dy
    assert np.isclose(mass, 0.1 * (np.pi * R * R / 2 - 2 * R * R / np.pi), rtol=2e-3)
    # u = u_amp * shape = h - h_ref here (u_amp = h_amp = 0.1), so hu = (h - 0.5) h; hv alike
    assert np.allclose(hu, (h - 0.5) * h, rtol=0, atol=1e-6) and np.array_equal(hu, hv)
    hf, huf, hvf, _, _ = I.bump(128, 128, W, W, bump_size=R)
    hc, huc, hvc, _, _ = I.bump(32, 32, W, W, bump_size=R, ref_nx=128, ref_ny=128)
    assert np.allclose(hc, I.downsample(hf, 4), rtol=0, atol=1e-6)
    assert np.allclose(huc, I.downsample(huf / hf, 4) * hc, rtol=0, atol=1e-6)
    assert h.min() >= 0.5 - 1e-6 and h.max() <= 0.6 + 1e-6


def test_shock_bubble_states():
    a = I.genShockBubble(160, 40, 1.4)
    rho, ru, E = a["rho"], a["rho_u"], a["E"]
    assert a["dx"] == 4.0 / 160 and a["dy"] == 1.0 / 40 and a["g"] == 0.0
    x = (np.arange(160) + 0.5) * a["dx"]
    left = x < 0.1
    assert np.allclose(rho[:, left], 3.8125) and np.allclose(ru[:, left], 3.8125 * 2.57669, rtol=1e-6)
    assert np.allclose(E[:, left], 0.5 * 3.8125 * 2.57669 ** 2 + 10.0 / 0.4, rtol=1e-6)
    bubble = np.isclose(rho, 0.1)
    assert abs(bubble.sum() * a["dx"] * a["dy"] - np.pi * 0.25 ** 2) < 0.01
    rest = ~left[None, :] & ~bubble
    assert np.allclose(rho[rest], 1.0) and np.allclose(E[rest], 1.0 / 0.4)
    assert a["boundary_conditions"].asCodedInt() == (3 << 24) | (3 << 16) | (2 << 8) | 2


def test_kelvin_helmholtz_zones():
    np.random.seed(3)
    a = I.genKelvinHelmholtz(64, 64, 1.4)
    rho, ru = a["rho"], a["rho_u"]
    assert set(np.unique(rho)) == {1.0, 2.0}
    assert np.allclose(ru[rho == 2.0], -1.0) and np.allclose(ru[rho == 1.0], 0.5)
    assert abs((rho == 2.0).mean() - 0.5) < 0.05  # the band y in (0.25, 0.75), up to the perturbation
    assert np.allclose(a["E"], 0.5 * rho * (ru / rho) ** 2 + 2.5 / 0.4)
    np.random.seed(3)
    b = I.genKelvinHelmholtz(64, 64, 1.4)
    assert np.array_equal(a["rho"], b["rho"])


@pytest.mark.parametrize("version", [0, 1])
def test_rayleigh_taylor_hydrostatic(version):
    a = I.genRayleighTaylor(32, 96, 1.4, version=version)
    rho, rv, E, g = a["rho"], a["rho_v"], a["E"], a["g"]
    y = (np.arange(96) + 0.5) * a["dy"] - 0.75
    v = rv / rho
    p = 0.4 * (E - 0.5 * rho * v * v)
    assert np.allclose(p, 2.5 - rho * g * y[:, None], rtol=0, atol=1e-5)
    assert rho[0].min() == 1.0 and rho[-1].max() == 2.0
    if version == 1:
        assert v.max() > 0.0
    else:
        assert not rv.any()
