"""SHMEMGrid shape rule and neighbour arithmetic, and the region-copy entry
point's argument checks (CPU only).

The grid shapes are worked by hand from the reference's rule
(SHMEMSimulatorGroup.py:97-156): factors of least sum over divisors d with
d*d < n, a lone 4 split into 2 x 2, padded with ones, sorted descending.  The
strict d*d < n is why 4 needs the special case and why 9 gives 9 x 1."""

import ctypes

import numpy as np
import pytest

from finitevolumegpu_amd import _lib
from finitevolumegpu_amd.SHMEMSimulator import SHMEMGrid

HAND = {1: [1, 1], 2: [2, 1], 3: [3, 1], 4: [2, 2], 5: [5, 1], 6: [3, 2], 7: [7, 1], 8: [4, 2], 9: [9, 1],
        10: [5, 2], 12: [4, 3], 16: [8, 2], 18: [6, 3], 24: [6, 4], 36: [9, 4]}


@pytest.mark.parametrize("n", sorted(HAND))
def test_grid_shape_matches_the_reference_rule(n):
    assert SHMEMGrid.getGrid(n, 2).tolist() == HAND[n]


def test_grid_shape_invariants():
    for n in range(1, 65):
        g = SHMEMGrid.getGrid(n, 2)
        assert int(np.prod(g)) == n and g[0] >= g[1] >= 1


def test_grid_rejects_non_integers():
    with pytest.raises(TypeError):
        SHMEMGrid.getGrid(4.0, 2)


class _Grid(SHMEMGrid):
    """The neighbour arithmetic without contexts."""

    def __init__(self, shape):
        self.grid = np.array(shape)


def test_neighbours_wrap_on_the_grid():
    g = _Grid([3, 2])
    # index = j * 3 + i
    assert g.getCoordinate(4) == (1, 1)
    assert g.getIndex(2, 1) == 5
    assert (g.getEast(2), g.getWest(0), g.getNorth(4), g.getSouth(1)) == (0, 2, 1, 4)
    for i in range(6):
        assert g.getWest(g.getEast(i)) == i and g.getSouth(g.getNorth(i)) == i


def _regions(*rs):
    arr = (_lib.RegionCopy * len(rs))()
    for k, (src, dst, sp, dp, w, h) in enumerate(rs):
        arr[k].src, arr[k].dst, arr[k].src_pitch, arr[k].dst_pitch, arr[k].width, arr[k].height = src, dst, sp, dp, w, h
    return arr


@pytest.mark.parametrize("region,msg", [
    ((0x1000, 0x2000, 256, 256, -1, 4), "negative"),
    ((None, 0x2000, 256, 256, 4, 4), "null"),
    ((0x1000, 0x2000, 8, 256, 4, 4), "pitch"),
])
def test_region_copy_validation(region, msg):
    lib = _lib.lib()
    assert lib.fvmi_copy_regions_f32(_regions(region), 1, None) < 0
    assert msg in lib.fvmi_last_error().decode()


def test_region_copy_of_nothing_launches_nothing():
    lib = _lib.lib()
    assert lib.fvmi_copy_regions_f32(None, 0, None) == 0
    # empty regions are skipped before any launch
    assert lib.fvmi_copy_regions_f32(_regions((0x1000, 0x2000, 256, 256, 0, 7)), 1, None) == 0
    assert lib.fvmi_copy_regions_f32(None, -1, None) < 0


@pytest.mark.parametrize("nx,ny,g", [(16, 12, 2), (5, 9, 1), (40, 24, 2)])
def test_exchange_extents(nx, ny, g):
    """Writes cover the ghost frame exactly once per leg (north/south: the ghost
    rows over the interior columns; east/west: the ghost columns over the full
    height); reads are interior cells, the neighbour's cells next to the shared
    edge; every region lies inside the padded array."""
    from finitevolumegpu_amd.SHMEMSimulator import exchange_extents
    ext = exchange_extents(nx, ny, g, g)
    H, W = ny + 2 * g, nx + 2 * g

    def mask(r):
        x, y, w, h = (int(v) for v in r)
        assert 0 <= x and 0 <= y and x + w <= W and y + h <= H
        m = np.zeros((H, W), dtype=int)
        m[y:y + h, x:x + w] = 1
        return m

    read_e, read_w, read_n, read_s = ext["read"]
    write_e, write_w, write_n, write_s = ext["write"]
    interior = np.zeros((H, W), dtype=int)
    interior[g:g + ny, g:g + nx] = 1
    ns = mask(write_n) + mask(write_s)
    assert ns.max() == 1 and (ns * interior).sum() == 0 and ns.sum() == 2 * g * nx
    ew = mask(write_e) + mask(write_w)
    assert ew.max() == 1 and (ew * interior).sum() == 0 and ew.sum() == 2 * g * H
    # reads: interior columns (all rows for east/west), interior rows for north/south
    assert (mask(read_n) * interior).sum() == g * nx and (mask(read_s) * interior).sum() == g * nx
    assert mask(read_e)[:, nx:nx + g].sum() == g * H and mask(read_w)[:, g:2 * g].sum() == g * H
    # a subdomain's north ghost rows receive its northern neighbour's southmost interior rows
    assert read_s[1] == g and write_n[1] == ny + g and read_n[1] == ny and write_s[1] == 0
