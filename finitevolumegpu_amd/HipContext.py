"""Device context, streams and events over libfvmi (replaces GPUSimulators/CudaContext.py).

``HipContext(device, context_flags, use_cache, autotuning)`` keeps the
constructor signature of ``CudaContext`` (CudaContext.py:45) and its public
attributes ``autotuner`` and ``synchronize()``.  There is no run-time kernel
compilation: the kernels are compiled ahead of time into libfvmi.so, so
``use_cache`` is accepted and ignored.
"""

import ctypes
import logging

from . import _lib


class Stream:
    """A non-blocking HIP stream (pycuda ``cuda.Stream``)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _lib.call("fvmi_stream_create", ctypes.byref(h))
        self.handle = h

    def synchronize(self):
        _lib.call("fvmi_stream_synchronize", self.handle)

    def wait_for_event(self, event):
        """Work queued after this call waits for ``event``."""
        _lib.call("fvmi_stream_wait_event", self.handle, event.handle)

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib._lib is not None:
            _lib.lib().fvmi_stream_destroy(self.handle)
            self.handle = None


class Event:
    """A HIP event (pycuda ``cuda.Event``): record / synchronize / time_since."""

    def __init__(self):
        h = ctypes.c_void_p()
        _lib.call("fvmi_event_create", ctypes.byref(h))
        self.handle = h

    def record(self, stream=None):
        _lib.call("fvmi_event_record", self.handle, stream.handle if stream is not None else None)
        return self

    def synchronize(self):
        _lib.call("fvmi_event_synchronize", self.handle)

    def time_since(self, start):
        """Milliseconds from ``start`` to this event."""
        ms = ctypes.c_float()
        _lib.call("fvmi_event_elapsed_ms", start.handle, self.handle, ctypes.byref(ms))
        return ms.value

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib._lib is not None:
            _lib.lib().fvmi_event_destroy(self.handle)
            self.handle = None


def device_count():
    n = ctypes.c_int()
    _lib.call("fvmi_device_count", ctypes.byref(n))
    return n.value


class HipContext:
    """GPU context for one device (CudaContext.py:45-120)."""

    def __init__(self, device=None, context_flags=None, use_cache=True, autotuning=True):
        self.logger = logging.getLogger(__name__)
        self.use_cache = use_cache
        if device is None:
            device = 0
        self.device = int(device)
        h = ctypes.c_void_p()
        _lib.call("fvmi_ctx_create", self.device, 0 if context_flags is None else int(context_flags), ctypes.byref(h))
        self.handle = h
        buf = ctypes.create_string_buffer(256)
        _lib.call("fvmi_device_name", self.device, buf, 256)
        self.device_name = buf.value.decode(errors="replace")
        cus = ctypes.c_int()
        _lib.call("fvmi_device_cu_count", self.device, ctypes.byref(cus))
        self.cu_count = cus.value
        free, total = self.mem_info()
        self.logger.info("Using device %d '%s', %d / %d MB free", self.device, self.device_name,
                         free // (1 << 20), total // (1 << 20))
        self.autotuner = None
        if autotuning:
            from .Autotuner import Autotuner
            self.autotuner = Autotuner()

    def make_current(self):
        _lib.call("fvmi_ctx_make_current", self.handle)

    def synchronize(self):
        _lib.call("fvmi_ctx_synchronize", self.handle)

    def mem_info(self):
        free = ctypes.c_size_t()
        total = ctypes.c_size_t()
        _lib.call("fvmi_mem_get_info", ctypes.byref(free), ctypes.byref(total))
        return free.value, total.value

    def __str__(self):
        return "HipContext device %d (%s)" % (self.device, self.device_name)

    def __del__(self):
        if getattr(self, "handle", None) is not None and _lib._lib is not None:
            _lib.lib().fvmi_ctx_destroy(self.handle)
            self.handle = None


# Drop-in name for code written against the reference.
CudaContext = HipContext
