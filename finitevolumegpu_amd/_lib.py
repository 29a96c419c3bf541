"""ctypes binding of libfvmi.so (declarations: include/fvmi.h).

There is no fallback: if the shared library is missing or cannot be loaded,
``lib()`` raises.  Build it with ``python -m finitevolumegpu_amd.build``.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfvmi.so")

c_int = ctypes.c_int
c_float = ctypes.c_float
c_size_t = ctypes.c_size_t
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p
P = ctypes.POINTER

FVMI_MATH_FAST = 0
FVMI_MATH_EXACT = 1
MEMCPY_HTOD = 1
MEMCPY_DTOH = 2
MEMCPY_DTOD = 3


class LaunchConfig(ctypes.Structure):
    """fvmi_launch_t"""
    _fields_ = [("tile_rows", c_int), ("waves_per_block", c_int), ("math", c_int), ("xcd_remap", c_int)]


class RegionCopy(ctypes.Structure):
    """fvmi_region_copy_t"""
    _fields_ = [("src", c_void_p), ("dst", c_void_p), ("src_pitch", c_int), ("dst_pitch", c_int),
                ("width", c_int), ("height", c_int)]


class Instance(ctypes.Structure):
    """fvmi_instance_t"""
    _fields_ = [("field", (c_void_p * 4) * 2), ("pitch", c_int), ("nx", c_int), ("ny", c_int), ("bc", c_int),
                ("dx", c_float), ("dy", c_float), ("g", c_float), ("theta", c_float),
                ("gamma", c_float), ("sqrt_g", c_float), ("inv_g", c_float), ("inv_dx", c_float),
                ("inv_dy", c_float), ("cfl", c_float * 2), ("dt_step", c_float),
                ("nt", c_int), ("active", c_int), ("fixed_dt", c_int),
                ("t", ctypes.c_double), ("t_end", ctypes.c_double), ("dt", ctypes.c_double),
                ("cfl_scale", ctypes.c_double)]


def _instance_dtype():
    """numpy structured dtype with the offsets of ``Instance`` (a view of a host
    table without per-field ctypes access)."""
    kinds = {c_int: "<i4", c_float: "<f4", ctypes.c_double: "<f8"}
    names, formats, offsets = [], [], []
    for name, ctype in Instance._fields_:
        if name == "field":
            fmt = ("<u8", (2, 4))
        elif name == "cfl":
            fmt = ("<f4", (2,))
        else:
            fmt = kinds[ctype]
        names.append(name)
        formats.append(fmt)
        offsets.append(getattr(Instance, name).offset)
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": ctypes.sizeof(Instance)})


INSTANCE_DTYPE = _instance_dtype()

SCHEME_IDS = {"kp07_dimsplit": 0, "hll2": 1, "waf": 2, "ee2d_kp07_dimsplit": 3}

_fp = P(c_float)

# name -> (restype, argtypes)
SIGNATURES = {
    "fvmi_last_error": (c_char_p, []),
    "fvmi_version": (c_int, []),
    "fvmi_device_count": (c_int, [P(c_int)]),
    "fvmi_ctx_create": (c_int, [c_int, c_int, P(c_void_p)]),
    "fvmi_ctx_destroy": (c_int, [c_void_p]),
    "fvmi_ctx_make_current": (c_int, [c_void_p]),
    "fvmi_ctx_synchronize": (c_int, [c_void_p]),
    "fvmi_ctx_device": (c_int, [c_void_p, P(c_int)]),
    "fvmi_device_name": (c_int, [c_int, c_char_p, c_int]),
    "fvmi_device_cu_count": (c_int, [c_int, P(c_int)]),
    "fvmi_mem_get_info": (c_int, [P(c_size_t), P(c_size_t)]),
    "fvmi_stream_create": (c_int, [P(c_void_p)]),
    "fvmi_stream_destroy": (c_int, [c_void_p]),
    "fvmi_stream_synchronize": (c_int, [c_void_p]),
    "fvmi_event_create": (c_int, [P(c_void_p)]),
    "fvmi_event_destroy": (c_int, [c_void_p]),
    "fvmi_event_record": (c_int, [c_void_p, c_void_p]),
    "fvmi_event_synchronize": (c_int, [c_void_p]),
    "fvmi_stream_wait_event": (c_int, [c_void_p, c_void_p]),
    "fvmi_event_elapsed_ms": (c_int, [c_void_p, c_void_p, P(c_float)]),
    "fvmi_malloc": (c_int, [P(c_void_p), c_size_t]),
    "fvmi_malloc_pitch": (c_int, [P(c_void_p), P(c_size_t), c_size_t, c_size_t]),
    "fvmi_free": (c_int, [c_void_p]),
    "fvmi_host_malloc": (c_int, [P(c_void_p), c_size_t]),
    "fvmi_host_free": (c_int, [c_void_p]),
    "fvmi_memcpy_async": (c_int, [c_void_p, c_void_p, c_size_t, c_int, c_void_p]),
    "fvmi_memcpy2d_async": (c_int, [c_void_p, c_size_t, c_void_p, c_size_t, c_size_t, c_size_t, c_int, c_void_p]),
    "fvmi_fill_f32": (c_int, [c_void_p, c_size_t, c_int, c_int, c_float, c_void_p]),
    "fvmi_copy_f32": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "fvmi_copy_regions_f32": (c_int, [c_void_p, c_int, c_void_p]),
    "fvmi_substep_plan": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_int, P(LaunchConfig), c_int,
                                  P(c_int), c_int, P(c_int)]),
    "fvmi_swe_kp07_dimsplit_substep": (c_int, [c_int, c_int, c_float, c_float, c_float, c_float, c_float, c_int, c_int]
                                       + [c_void_p, c_int] * 6 + [c_void_p, P(LaunchConfig), c_void_p]),
    "fvmi_swe_hll2_substep": (c_int, [c_int, c_int, c_float, c_float, c_float, c_float, c_float, c_int, c_int]
                              + [c_void_p, c_int] * 6 + [c_void_p, P(LaunchConfig), c_void_p]),
    "fvmi_swe_waf_substep": (c_int, [c_int, c_int, c_float, c_float, c_float, c_float, c_int, c_int]
                             + [c_void_p, c_int] * 6 + [c_void_p, P(LaunchConfig), c_void_p]),
    "fvmi_euler_kp07_dimsplit_substep": (c_int, [c_int, c_int, c_float, c_float, c_float, c_float, c_float, c_float,
                                                 c_int, c_int] + [c_void_p, c_int] * 8
                                         + [c_void_p, c_int, c_int, c_int, c_int, P(LaunchConfig), c_void_p]),
    "fvmi_swe_lxf_substep": (c_int, [c_int, c_int, c_float, c_float, c_float, c_float, c_int]
                             + [c_void_p, c_int] * 6 + [c_void_p, c_int, P(LaunchConfig), c_void_p]),
    "fvmi_swe_hll_substep": (c_int, [c_int, c_int, c_float, c_float, c_float, c_float, c_int]
                             + [c_void_p, c_int] * 6 + [c_void_p, c_int, P(LaunchConfig), c_void_p]),
    "fvmi_swe_force_substep": (c_int, [c_int, c_int, c_float, c_float, c_float, c_float, c_int]
                               + [c_void_p, c_int] * 6 + [c_void_p, c_int, P(LaunchConfig), c_void_p]),
    "fvmi_swe_kp07_substep": (c_int, [c_int, c_int, c_float, c_float, c_float, c_float, c_float, c_int, c_int]
                              + [c_void_p, c_int] * 6 + [c_void_p, c_int, P(LaunchConfig), c_void_p]),
    "fvmi_halo_copy_periodic_x": (c_int, [c_int, c_int, c_int, P(c_void_p), P(c_int), c_void_p]),
    "fvmi_halo_exchange_periodic": (c_int, [c_int, c_int, c_int, P(c_void_p), P(c_int), c_int, c_int, c_void_p]),
    "fvmi_field_stats": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "fvmi_ensemble_prepare": (c_int, [c_void_p, c_int, c_void_p]),
    "fvmi_ensemble_substep": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, P(LaunchConfig), c_void_p]),
    "fvmi_ensemble_advance": (c_int, [c_void_p, c_int, c_void_p]),
    "fvmi_ensemble_count_active": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "fvmi_stream_begin_capture": (c_int, [c_void_p]),
    "fvmi_stream_end_capture": (c_int, [c_void_p, P(c_void_p)]),
    "fvmi_graph_launch": (c_int, [c_void_p, c_void_p]),
    "fvmi_graph_destroy": (c_int, [c_void_p]),
}

_lib = None


class FvmiError(RuntimeError):
    """A libfvmi call returned a non-zero status."""


def lib():
    """Loads libfvmi.so once and declares every exported signature."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FvmiError(
                "libfvmi.so is not built (expected at %s); run `python -m finitevolumegpu_amd.build`" % LIB_PATH)
        handle = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc, what=""):
    """Raises FvmiError with the library's message when rc != 0."""
    if rc != 0:
        msg = lib().fvmi_last_error().decode(errors="replace")
        raise FvmiError("%s failed (%d): %s" % (what, rc, msg))
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args), name)
