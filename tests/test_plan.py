"""The substep launch plan, checked on the host through fvmi_substep_plan
(the same planner and segment mapping the kernels run; no device call): for
every scheme, domain, sub-area, launch shape and CU count below, the waves'
segments cover every row of every strip of the sub-area exactly once (124
columns for the schemes that march two columns per lane, 60 or 62 for the
others).
Also the choice between the round-weighted and the equal-segment plans and the
wave budget of the equal-segment plan."""

import ctypes

import numpy as np
import pytest

from finitevolumegpu_amd import _lib

SCHEMES = {"kp07": 0, "hll2": 1, "waf": 2, "ee2d": 3, "lxf": 4, "hll": 5, "force": 6, "kp07u": 7}
# output columns per strip: 64 lanes less the ghost lanes, times the columns per lane
STRIP = {"kp07": 60, "hll2": 60, "ee2d": 124, "ee2d_exact": 60, "waf": 60, "lxf": 62, "hll": 62, "force": 62,
         "kp07u": 60, "kp07_exact": 60}


def plan(scheme, nx, ny, sub=(0, 0, 0, 0), tile_rows=0, waves_per_block=4, xcd_remap=True, cu_count=256):
    lib = _lib.lib()
    math = _lib.FVMI_MATH_EXACT if scheme.endswith("_exact") else _lib.FVMI_MATH_FAST
    scheme = scheme.replace("_exact", "")
    launch = _lib.LaunchConfig(tile_rows, waves_per_block, math, 1 if xcd_remap else 2)
    n = ctypes.c_int()
    rc = lib.fvmi_substep_plan(SCHEMES[scheme], nx, ny, *sub, ctypes.byref(launch), cu_count, None, 0,
                               ctypes.byref(n))
    assert rc != 0 and n.value > 0  # sizes the buffer first
    segs = np.zeros((n.value, 3), dtype=np.int32)
    rc = lib.fvmi_substep_plan(SCHEMES[scheme], nx, ny, *sub, ctypes.byref(launch), cu_count,
                               segs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n.value, ctypes.byref(n))
    assert rc == 0, lib.fvmi_last_error().decode()
    return segs


def assert_covers(segs, width, height, strip=60):
    strips = -(-width // strip)
    cover = np.zeros((strips, height), dtype=np.int32)
    for s, r, c in segs:
        if s < 0:
            assert r == 0 and c == 0
            continue
        assert 0 <= s < strips and c >= 1 and 0 <= r and r + c <= height, (s, r, c)
        cover[s, r:r + c] += 1
    assert (cover == 1).all(), np.argwhere(cover != 1)[:5]


CASES = [
    ("kp07", 4096, 4096, (0, 0, 0, 0), 0, 4, True, 256),
    ("ee2d", 8192, 8192, (0, 0, 0, 0), 0, 4, True, 256),
    ("hll2", 3000, 2500, (0, 0, 0, 0), 0, 4, True, 256),
    ("waf", 4096, 4096, (0, 0, 0, 0), 0, 4, True, 256),
    ("kp07", 1000, 4000, (0, 0, 0, 0), 0, 4, True, 256),
    ("kp07", 6000, 777, (0, 0, 0, 0), 0, 4, True, 256),
    ("ee2d", 512, 8192, (2, 0, 510, 8192), 0, 4, True, 256),
    ("ee2d", 700, 900, (2, 2, 698, 898), 0, 4, True, 304),
    ("ee2d_exact", 8192, 8192, (0, 0, 0, 0), 0, 4, True, 256),
    ("ee2d_exact", 700, 900, (2, 2, 698, 898), 0, 4, True, 304),
    ("ee2d", 701, 900, (699, 0, 701, 900), 0, 4, True, 256),
    ("hll2", 127, 60, (0, 0, 0, 0), 0, 4, True, 256),
    ("kp07", 257, 129, (0, 0, 0, 0), 0, 4, True, 256),
    ("kp07", 13, 11, (0, 0, 0, 0), 0, 4, True, 256),
    ("kp07", 4096, 4096, (0, 0, 0, 0), 70, 4, True, 256),
    ("kp07", 4096, 4096, (0, 0, 0, 0), 0, 1, True, 256),
    ("hll2", 4096, 4096, (0, 0, 0, 0), 0, 2, False, 256),
    ("kp07", 20000, 300, (0, 0, 0, 0), 0, 4, True, 256),
    ("kp07", 2048, 2048, (0, 0, 0, 0), 0, 4, True, 80),
    ("lxf", 4096, 4096, (0, 0, 0, 0), 0, 4, True, 256),
    ("hll", 3000, 2500, (0, 0, 0, 0), 0, 4, True, 256),
    ("force", 777, 333, (0, 0, 0, 0), 0, 4, True, 256),
    ("kp07u", 4096, 4096, (0, 0, 0, 0), 0, 4, True, 256),
    ("kp07u", 4096, 4096, (0, 0, 0, 0), 16, 2, True, 256),
    ("lxf", 13, 11, (0, 0, 0, 0), 0, 4, True, 256),
]


@pytest.mark.parametrize("scheme,nx,ny,sub,tr,wpb,remap,cus", CASES)
def test_segments_cover_the_sub_area_once(scheme, nx, ny, sub, tr, wpb, remap, cus):
    segs = plan(scheme, nx, ny, sub, tr, wpb, remap, cus)
    x0, y0, x1, y1 = sub
    x1 = x1 or nx
    y1 = y1 or ny
    assert_covers(segs, x1 - x0, y1 - y0, STRIP[scheme])


def test_round_weighted_plan_at_4096():
    segs = plan("kp07", 4096, 4096)
    assert len(segs) == 4 * 256 * 4  # 4 rounds of one 4-wave workgroup per CU
    active = segs[segs[:, 0] >= 0]
    # 69 strips of 60 columns in 18 bands of 4 (the last band three quarters
    # full); 256 CUs / 18 bands = 14 row groups per band, 4 parts each
    assert len(active) == 69 * 14 * 4
    # workgroup p is in round (p // 8) // 32; earlier rounds take longer parts (14:11:9:6)
    blocks = np.arange(len(segs)) // 4
    rounds = (blocks // 8) // 32
    mean = [segs[(rounds == k) & (segs[:, 0] >= 0), 2].mean() for k in range(4)]
    assert mean[0] > mean[1] > mean[2] > mean[3]
    # a 293-row group (4096 rows in 14 groups) cut 14:11:9:6, the last part taking the remainder
    first = sorted((r, c) for s, r, c in active if s == 0 and r < 293)
    assert [c for _, c in first] == [293 * 14 // 40, 293 * 11 // 40, 293 * 9 // 40,
                                     293 - 293 * 14 // 40 - 293 * 11 // 40 - 293 * 9 // 40] == [102, 80, 65, 46]
    # the 4 waves of a workgroup take the same rows of 4 adjacent strips
    for b in range(0, len(segs), 4):
        grp = segs[b:b + 4]
        live = grp[grp[:, 0] >= 0]
        if len(live) > 1:
            assert (live[:, 1] == live[0, 1]).all() and (live[:, 2] == live[0, 2]).all()
            assert (np.diff(live[:, 0]) == 1).all() and live[0, 0] % 4 == 0


def test_euler_fast_plan_uses_two_column_strips():
    """The fast Euler kernel marches two columns per lane: 8192 columns in 67
    strips of 124 (the exact one in 137 of 60)."""
    fast = plan("ee2d", 8192, 8192)
    exact = plan("ee2d_exact", 8192, 8192)
    assert fast[:, 0].max() == 66 and exact[:, 0].max() == 136


def test_waf_keeps_equal_segments():
    segs = plan("waf", 4096, 4096)
    counts = segs[segs[:, 0] >= 0, 2]
    assert len(np.unique(counts[:-1])) <= 2  # equal segments (the last of a strip shorter)


def test_equal_plan_stays_within_four_waves_per_simd():
    # 137 strips at 8192 columns: the equal plan must not exceed 4 waves per SIMD
    segs = plan("ee2d_exact", 8192, 8192, tile_rows=0, waves_per_block=2)
    assert (segs[:, 0] >= 0).sum() <= 256 * 4 * 4


def test_unsplit_round_plan_for_hll_and_kp07_only():
    for key, strips in (("hll", 67), ("kp07u", 69)):
        segs = plan(key, 4096, 4096)
        assert len(segs) == 4 * 256 * 4
        assert (segs[:, 0] >= 0).sum() == strips * (1024 // strips) * 4
    for key in ("lxf", "force"):
        counts = plan(key, 4096, 4096)[:, 2]
        assert len(np.unique(counts[counts > 0])) <= 2  # equal segments (the last of a strip shorter)


def test_plan_rejects_bad_arguments():
    lib = _lib.lib()
    n = ctypes.c_int()
    assert lib.fvmi_substep_plan(8, 64, 64, 0, 0, 0, 0, None, 256, None, 0, ctypes.byref(n)) < 0
    assert lib.fvmi_substep_plan(4, 64, 64, 1, 0, 0, 0, None, 256, None, 0, ctypes.byref(n)) < 0
    assert lib.fvmi_substep_plan(0, 64, 64, 0, 0, 65, 0, None, 256, None, 0, ctypes.byref(n)) < 0
