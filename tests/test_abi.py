"""C ABI of libfvmi.so (CPU only: no call reaches a GPU).

* the library loads and exports every function include/fvmi.h declares;
* the ctypes binding declares exactly those functions;
* argument validation fails loudly, before any device work, with the
  reference's error behaviour (Neumann boundaries rejected, Simulator.py:68-72).
"""

import ctypes
import os
import re

import pytest

from finitevolumegpu_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "fvmi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fvmi_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = header_functions()
    for required in ("fvmi_swe_kp07_dimsplit_substep", "fvmi_swe_hll2_substep", "fvmi_swe_waf_substep",
                     "fvmi_euler_kp07_dimsplit_substep", "fvmi_halo_copy_periodic_x", "fvmi_ctx_create",
                     "fvmi_malloc_pitch", "fvmi_memcpy2d_async", "fvmi_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_matches_header():
    assert sorted(_lib.SIGNATURES) == header_functions()


def test_version_without_gpu():
    assert _lib.lib().fvmi_version() == 1


def _substep(nx=16, ny=16, bc=(3 << 24) | (3 << 16) | (3 << 8) | 3, step=0, dt=0.1, ptr=0x1000, pitch=256):
    lib = _lib.lib()
    return lib.fvmi_swe_kp07_dimsplit_substep(nx, ny, 1.0, 1.0, dt, 9.81, 1.3, step, bc,
                                              ptr, pitch, ptr + 0x100000, pitch, ptr + 0x200000, pitch,
                                              ptr + 0x300000, pitch, ptr + 0x400000, pitch, ptr + 0x500000, pitch,
                                              None, None, None)


def _err():
    return _lib.lib().fvmi_last_error().decode()


def test_validation_rejects_tiny_grid():
    assert _substep(nx=1) == -1
    assert "too small" in _err()


def test_validation_rejects_neumann():
    neumann = (1 << 24) | (3 << 16) | (3 << 8) | 3
    assert _substep(bc=neumann) == -3
    assert "Neumann" in _err()


def test_validation_rejects_bad_step_dt_pitch():
    assert _substep(step=2) == -1
    assert _substep(dt=float("nan")) == -1
    assert _substep(pitch=16) == -1
    assert "pitch" in _err()


def test_validation_rejects_fields_of_2gib():
    # masked lanes store at a 2 GiB offset, so every field must stay below it
    assert _substep(pitch=((1 << 31) // 20 + 4) & ~3) == -1  # (16 + 4) rows
    assert "2 GiB" in _err()
    assert _first_order("fvmi_swe_lxf_substep", pitch=((1 << 31) // 18 + 4) & ~3) == -1  # (16 + 2) rows
    assert "2 GiB" in _err()


def test_validation_rejects_aliasing():
    lib = _lib.lib()
    p = 0x1000
    rc = lib.fvmi_swe_hll2_substep(8, 8, 1.0, 1.0, 0.1, 9.81, 1.8, 0, 0, p, 64, p + 4096, 64, p + 8192, 64,
                                   p, 64, p + 4096 * 4, 64, p + 4096 * 5, 64, None, None, None)
    assert rc == -1 and "alias" in _err()


def test_validation_rejects_bad_subarea():
    lib = _lib.lib()
    p = 0x1000
    args = [p + 0x10000 * k for k in range(8)]
    rc = lib.fvmi_euler_kp07_dimsplit_substep(8, 8, 1.0, 1.0, 0.1, 0.1, 1.4, 1.3, 0, 0,
                                              args[0], 64, args[1], 64, args[2], 64, args[3], 64,
                                              args[4], 64, args[5], 64, args[6], 64, args[7], 64,
                                              None, 4, 0, 2, 8, None, None)
    assert rc == -1 and "sub-area" in _err()


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(_lib.FvmiError):
        _lib.lib()


def _first_order(fn, nx=16, ny=16, dt=0.1, slot=0, pitch=256, bc=0):
    p = 0x1000
    return getattr(_lib.lib(), fn)(nx, ny, 1.0, 1.0, dt, 9.81, bc, p, pitch, p + 0x100000, pitch, p + 0x200000, pitch,
                                   p + 0x300000, pitch, p + 0x400000, pitch, p + 0x500000, pitch, None, slot, None,
                                   None)


@pytest.mark.parametrize("fn", ["fvmi_swe_lxf_substep", "fvmi_swe_hll_substep", "fvmi_swe_force_substep"])
def test_first_order_validation(fn):
    assert _first_order(fn, nx=1) == -1 and "too small" in _err()
    assert _first_order(fn, slot=2) == -1 and "cfl_slot" in _err()
    assert _first_order(fn, pitch=64) == -1 and "pitch" in _err()  # (16 + 2) * 4 = 72 bytes needed
    assert _first_order(fn, nx=40000, ny=40000) == -1 and "too large" in _err()
    neumann = (1 << 24) | (3 << 16) | (3 << 8) | 3
    assert _first_order(fn, bc=neumann) == -3
    if fn != "fvmi_swe_hll_substep":  # the centred fluxes divide by dt
        assert _first_order(fn, dt=0.0) == -1 and "dt" in _err()


def test_kp07_step_order_validation():
    lib = _lib.lib()
    p = 0x1000
    args = [p + 0x100000 * k for k in range(6)]

    def call(step_order):
        return lib.fvmi_swe_kp07_substep(16, 16, 1.0, 1.0, 0.1, 9.81, 1.3, step_order, 0, args[0], 256, args[1], 256,
                                         args[2], 256, args[3], 256, args[4], 256, args[5], 256, None, 0, None, None)
    assert call((2 << 16) | 2) == -1 and "order" in _err()
    assert call((0 << 16) | 3) == -1
    assert call((1 << 16) | 1) == -1


def test_ensemble_substep_validation():
    lib = _lib.lib()
    assert lib.fvmi_ensemble_substep(0, None, None, None, 0, 0, None, None) == -1
    assert lib.fvmi_ensemble_substep(7, 0x1000, 0x2000, 0x3000, 0, 0, None, None) == -1 and "scheme" in _err()
    assert lib.fvmi_ensemble_substep(0, 0x1000, 0x2000, 0x3000, 0, 2, None, None) == -1 and "step" in _err()


def test_field_stats_validation():
    """The pitch must hold the interior and both ghost frames of the stated
    widths (a one-ghost-cell field of width 126 has pitch 512 >= 128 * 4 but
    not (126 + 4) * 4)."""
    lib = _lib.lib()
    f = ctypes.c_void_p(0x1000)
    out = ctypes.c_void_p(0x2000)
    assert lib.fvmi_field_stats(126, 8, 1, 1, f, 511, out, out, None) < 0
    assert "field_stats" in _err()
    assert lib.fvmi_field_stats(126, 8, -1, 1, f, 1024, out, out, None) < 0
    assert lib.fvmi_field_stats(0, 8, 1, 1, f, 1024, out, out, None) < 0
    assert lib.fvmi_field_stats(126, 8, 1, 1, None, 1024, out, out, None) < 0


def test_fast_euler_needs_8_byte_aligned_fields():
    """The fast Euler kernel moves two columns per lane with 8-byte accesses:
    a field pointer or pitch that is only 4-byte aligned is rejected before any
    device work (no call here passes validation: the pointers are not device
    memory)."""
    lib = _lib.lib()
    base = [0x100000 + 0x10000 * k for k in range(8)]

    def euler(ptrs, pitch, math):
        launch = _lib.LaunchConfig(0, 4, math, 1)
        args = []
        for p in ptrs:
            args += [p, pitch]
        return lib.fvmi_euler_kp07_dimsplit_substep(8, 8, 1.0, 1.0, 0.1, 0.1, 1.4, 1.3, 0, 0, *args, None,
                                                    0, 0, 0, 0, ctypes.byref(launch), None)

    odd_ptr = [base[0] + 4] + base[1:]
    assert euler(odd_ptr, 64, _lib.FVMI_MATH_FAST) == -1 and "aligned" in _err()
    assert euler(base, 52, _lib.FVMI_MATH_FAST) == -1 and "multiple of 8" in _err()
